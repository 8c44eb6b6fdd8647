// Host symbolic analysis for the supernodal Cholesky (see symbolic.h).
//
// The ordering replaces the reference backends' own orderings (CSparse AMD in ManagedSpdSolver.cs:31, METIS nested
// dissection inside MKL PARDISO with iparm[1] = 3 in MklPardisoSolver.cs:23): a nested dissection of the block graph
// whose separators are level sets of d(a, v) - d(b, v) between the two ends a, b of a pseudo-diameter. Everything
// after the ordering (elimination tree, column structure, fill) is a deterministic function of (pattern, ordering)
// and is what the tests compare bit-exactly against the oracle's brute-force symbolic elimination.
#include "symbolic.h"

#include <algorithm>
#include <cstring>
#include <numeric>

namespace ifea {
namespace {

struct Graph {
  int n = 0;
  std::vector<int> ptr, adj;  // no self loops
  std::vector<int> w;         // vertex weights (block sizes)
};

// ------------------------------------------------------------------------------------------- nested dissection
struct ND {
  const Graph& g;
  int leaf;
  std::vector<int> mark, da, db, queue;
  int label = 0;
  std::vector<int> out;
  explicit ND(const Graph& gr, int leaf_) : g(gr), leaf(leaf_) {
    mark.assign(g.n, -1);
    da.assign(g.n, -1);
    db.assign(g.n, -1);
    queue.reserve(g.n);
  }

  // BFS restricted to vertices with mark == lab; returns the last vertex reached (max distance, min degree tie).
  int bfs(int src, int lab, std::vector<int>& dist, std::vector<int>* reached) {
    queue.clear();
    queue.push_back(src);
    dist[src] = 0;
    // use a per-call stamp: reset only reached vertices afterwards
    size_t h = 0;
    int last = src;
    while (h < queue.size()) {
      int v = queue[h++];
      last = v;
      for (int p = g.ptr[v]; p < g.ptr[v + 1]; p++) {
        int u = g.adj[p];
        if (mark[u] != lab || dist[u] >= 0) continue;
        dist[u] = dist[v] + 1;
        queue.push_back(u);
      }
    }
    // farthest with the smallest degree
    int dmax = dist[last];
    int best = last, bdeg = g.ptr[last + 1] - g.ptr[last];
    for (size_t q = queue.size(); q-- > 0;) {
      int v = queue[q];
      if (dist[v] != dmax) break;
      int dg = g.ptr[v + 1] - g.ptr[v];
      if (dg < bdeg || (dg == bdeg && v < best)) { best = v; bdeg = dg; }
    }
    if (reached) *reached = queue;
    return best;
  }

  void order(std::vector<int> verts) {
    // explicit stack of pending subsets: each entry is ordered after its children, so recurse manually
    recurse(verts);
  }

  void recurse(std::vector<int>& V) {
    if ((int)V.size() <= leaf) {
      std::sort(V.begin(), V.end());
      out.insert(out.end(), V.begin(), V.end());
      return;
    }
    int lab = label++;
    for (int v : V) { mark[v] = lab; da[v] = -1; db[v] = -1; }
    // connected components
    std::vector<int> comp;
    bfs(V[0], lab, da, &comp);
    if (comp.size() < V.size()) {
      std::vector<int> rest;
      rest.reserve(V.size() - comp.size());
      for (int v : V) if (da[v] < 0) rest.push_back(v);
      for (int v : V) mark[v] = -1;
      std::vector<int> c1 = comp;
      recurse(c1);
      recurse(rest);
      return;
    }
    for (int v : V) da[v] = -1;
    int a = bfs(V[0], lab, da, nullptr);
    for (int v : V) da[v] = -1;
    int b = bfs(a, lab, da, nullptr);
    for (int v : V) db[v] = -1;
    bfs(b, lab, db, nullptr);
    // f = da - db in [-D, D]
    int D = da[b];
    std::vector<long long> hist(2 * D + 3, 0);
    long long W = 0;
    for (int v : V) { hist[da[v] - db[v] + D] += g.w[v]; W += g.w[v]; }
    // choose c: separator {f == c, c+1}, A = {f < c}, B = {f > c+1}
    long long best = -1, bestA = 0;
    int bc = 0;
    long long below = 0;
    for (int c = 0; c + 1 < (int)hist.size(); c++) {
      long long sep = hist[c] + hist[c + 1];
      long long A = below, B = W - below - sep;
      below += hist[c];
      if (std::min(A, B) < (long long)(0.3 * W)) continue;
      // prefer small separators, then balance
      if (best < 0 || sep < best || (sep == best && std::llabs(A - B) < std::llabs(bestA - (W - bestA - best)))) {
        best = sep; bc = c; bestA = A;
      }
    }
    std::vector<int> A, B, S;
    if (best >= 0) {
      for (int v : V) {
        int f = da[v] - db[v] + D;
        if (f < bc) A.push_back(v);
        else if (f > bc + 1) B.push_back(v);
        else S.push_back(v);
      }
    } else {
      // fallback: a BFS level from a (every edge joins equal or adjacent levels)
      int L = 0;
      for (int v : V) L = std::max(L, da[v]);
      std::vector<long long> lw(L + 1, 0);
      for (int v : V) lw[da[v]] += g.w[v];
      long long acc = 0;
      int m = 0;
      for (; m <= L; m++) { if (acc + lw[m] > W / 2) break; acc += lw[m]; }
      m = std::min(m, L);
      for (int v : V) {
        if (da[v] < m) A.push_back(v);
        else if (da[v] > m) B.push_back(v);
        else S.push_back(v);
      }
      if (A.empty() && B.empty()) {
        std::sort(V.begin(), V.end());
        out.insert(out.end(), V.begin(), V.end());
        for (int v : V) mark[v] = -1;
        return;
      }
    }
    for (int v : V) mark[v] = -1;
    std::vector<int>().swap(V);
    recurse(A);
    recurse(B);
    std::sort(S.begin(), S.end());
    out.insert(out.end(), S.begin(), S.end());
  }
};

// Geometric nested dissection: cut the subset along its longest axis near the median (in a wide coordinate gap of the
// middle tenth); the separator is the part of the upper half adjacent to the lower half (every edge between the halves
// ends in it).
struct GeoND {
  const Graph& g;
  const std::vector<double>& xyz;  // 3 per block
  int leaf;
  std::vector<int> mark;
  std::vector<int> out;
  GeoND(const Graph& gr, const std::vector<double>& c, int leaf_) : g(gr), xyz(c), leaf(leaf_) { mark.assign(g.n, 0); }
  void recurse(std::vector<int>& V) {
    if ((int)V.size() <= leaf) {
      std::sort(V.begin(), V.end());
      out.insert(out.end(), V.begin(), V.end());
      return;
    }
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int v : V)
      for (int a = 0; a < 3; a++) { lo[a] = std::min(lo[a], xyz[3 * v + a]); hi[a] = std::max(hi[a], xyz[3 * v + a]); }
    int ax = 0;
    for (int a = 1; a < 3; a++) if (hi[a] - lo[a] > hi[ax] - lo[ax]) ax = a;
    std::vector<std::pair<double, int>> key(V.size());
    for (size_t q = 0; q < V.size(); q++) key[q] = {xyz[3 * V[q] + ax], V[q]};
    std::sort(key.begin(), key.end());
    // cut in a gap between consecutive coordinates near the median: on a layered mesh this falls between two node
    // layers, so the separator is the coupling depth (two layers for the two-ring pattern) rather than a split layer
    // plus that depth. Among the gaps of the middle tenth at least half as wide as the widest, the one nearest the
    // median is taken.
    const size_t n = V.size();
    size_t mid = n / 2;
    const size_t q0 = std::max<size_t>(1, (9 * n) / 20), q1 = std::min(n - 1, (11 * n) / 20);
    double widest = 0.0;
    for (size_t q = q0; q <= q1; q++) widest = std::max(widest, key[q].first - key[q - 1].first);
    size_t bestd = n;
    for (size_t q = q0; q <= q1; q++) {
      const size_t d = q > n / 2 ? q - n / 2 : n / 2 - q;
      if (key[q].first - key[q - 1].first >= 0.5 * widest && d < bestd) { bestd = d; mid = q; }
    }
    // A = the elements below the cut, B = the rest
    for (size_t q = 0; q < V.size(); q++) mark[key[q].second] = q < mid ? 1 : 2;
    std::vector<int> A, B, S;
    for (size_t q = 0; q < V.size(); q++) {
      int v = key[q].second;
      if (mark[v] == 1) { A.push_back(v); continue; }
      bool sep = false;
      for (int p = g.ptr[v]; p < g.ptr[v + 1] && !sep; p++) sep = mark[g.adj[p]] == 1;
      (sep ? S : B).push_back(v);
    }
    for (int v : V) mark[v] = 0;
    if (A.empty() || B.empty()) {
      std::sort(V.begin(), V.end());
      out.insert(out.end(), V.begin(), V.end());
      return;
    }
    std::vector<int>().swap(V);
    recurse(A);
    recurse(B);
    std::sort(S.begin(), S.end());
    out.insert(out.end(), S.begin(), S.end());
  }
};

// Elimination tree (Liu, path compression) of a symmetric graph given in the new numbering.
void etree_graph(int n, const std::vector<int>& ptr, const std::vector<int>& adj, const std::vector<int>& pos,
                 const std::vector<int>& order, std::vector<int>& parent) {
  parent.assign(n, -1);
  std::vector<int> anc(n, -1);
  for (int J = 0; J < n; J++) {
    int v = order[J];
    for (int p = ptr[v]; p < ptr[v + 1]; p++) {
      int i = pos[adj[p]];
      while (i != -1 && i < J) {
        int nx = anc[i];
        anc[i] = J;
        if (nx == -1) { parent[i] = J; break; }
        i = nx;
      }
    }
  }
}

void postorder(int n, const std::vector<int>& parent, std::vector<int>& post) {
  std::vector<int> head(n, -1), next(n, -1);
  for (int j = n - 1; j >= 0; j--) {
    if (parent[j] == -1) continue;
    next[j] = head[parent[j]];
    head[parent[j]] = j;
  }
  post.clear();
  post.reserve(n);
  std::vector<int> stack;
  for (int j = 0; j < n; j++) {
    if (parent[j] != -1) continue;
    stack.push_back(j);
    while (!stack.empty()) {
      int p = stack.back();
      int i = head[p];
      if (i == -1) { stack.pop_back(); post.push_back(p); }
      else { head[p] = next[i]; stack.push_back(i); }
    }
  }
}

}  // namespace

int symbolic_analyze(int n, const int* colPtr, const int* rowIdx, const int* userPerm, const SymbolicOptions& opt,
                     Symbolic& S) {
  S = Symbolic();
  S.n = n;
  if (n <= 0 || !colPtr || !rowIdx) { S.error = "empty pattern"; return -1; }
  for (int j = 0; j < n; j++) {
    if (colPtr[j + 1] < colPtr[j]) { S.error = "colPtr not monotone"; return -1; }
    bool diag = false;
    for (int p = colPtr[j]; p < colPtr[j + 1]; p++) {
      int i = rowIdx[p];
      if (i < 0 || i >= n) { S.error = "row index out of range"; return -1; }
      if (p > colPtr[j] && rowIdx[p - 1] >= i) { S.error = "row indices not sorted"; return -1; }
      if (i == j) diag = true;
    }
    if (!diag) { S.error = "structurally missing diagonal"; return -1; }
  }
  std::vector<int> up;
  if (userPerm) {
    up.assign(userPerm, userPerm + n);
    std::vector<char> seen(n, 0);
    for (int k = 0; k < n; k++) {
      if (up[k] < 0 || up[k] >= n || seen[up[k]]) { S.error = "invalid permutation"; return -2; }
      seen[up[k]] = 1;
    }
  }
  // ---- 1. block compression (consecutive identical columns)
  std::vector<int> blk_of(n), bfirst;
  for (int j = 0; j < n; j++) {
    bool same = j > 0 && (colPtr[j + 1] - colPtr[j]) == (colPtr[j] - colPtr[j - 1]) &&
                std::memcmp(rowIdx + colPtr[j], rowIdx + colPtr[j - 1], sizeof(int) * (colPtr[j + 1] - colPtr[j])) == 0;
    if (!same) bfirst.push_back(j);
    blk_of[j] = (int)bfirst.size() - 1;
  }
  int nb = (int)bfirst.size();
  bfirst.push_back(n);
  // a caller permutation must keep every block contiguous and in order, else work on scalar columns
  if (userPerm) {
    bool aligned = true;
    for (int k = 0; k < n && aligned; k++) {
      int c = up[k], b = blk_of[c];
      if (c == bfirst[b]) {
        for (int t = 1; t < bfirst[b + 1] - bfirst[b]; t++)
          if (k + t >= n || up[k + t] != c + t) { aligned = false; break; }
      } else if (k == 0 || up[k - 1] != c - 1) {
        aligned = false;
      }
    }
    if (!aligned) {
      bfirst.resize(n + 1);
      for (int j = 0; j <= n; j++) bfirst[j] = j;
      for (int j = 0; j < n; j++) blk_of[j] = j;
      nb = n;
    }
  }
  Graph g;
  g.n = nb;
  g.ptr.assign(nb + 1, 0);
  g.w.resize(nb);
  for (int b = 0; b < nb; b++) {
    g.w[b] = bfirst[b + 1] - bfirst[b];
    int j = bfirst[b];
    int last = -1;
    for (int p = colPtr[j]; p < colPtr[j + 1]; p++) {
      int ob = blk_of[rowIdx[p]];
      if (ob == b || ob == last) continue;
      g.adj.push_back(ob);
      last = ob;
    }
    g.ptr[b + 1] = (int)g.adj.size();
  }
  // symmetric check (the pattern must be structurally symmetric)
  // ---- 2. ordering of blocks
  std::vector<int> order;
  if (userPerm) {
    for (int k = 0; k < n; k++) if (up[k] == bfirst[blk_of[up[k]]]) order.push_back(blk_of[up[k]]);
  } else if (opt.coords) {
    std::vector<double> bxyz(3 * (size_t)nb);
    for (int b = 0; b < nb; b++)
      for (int a = 0; a < 3; a++) bxyz[3 * b + a] = opt.coords[3 * (size_t)bfirst[b] + a];
    GeoND nd(g, bxyz, std::max(1, opt.leaf_blocks));
    std::vector<int> all(nb);
    std::iota(all.begin(), all.end(), 0);
    nd.recurse(all);
    order.swap(nd.out);
  } else {
    ND nd(g, std::max(1, opt.leaf_blocks));
    std::vector<int> all(nb);
    std::iota(all.begin(), all.end(), 0);
    nd.order(all);
    order.swap(nd.out);
  }
  if ((int)order.size() != nb) { S.error = "ordering lost blocks"; return -2; }
  std::vector<int> pos(nb);
  for (int J = 0; J < nb; J++) pos[order[J]] = J;
  // ---- 3. etree, postorder (composed into the order), etree again
  std::vector<int> bparent, post;
  etree_graph(nb, g.ptr, g.adj, pos, order, bparent);
  if (!userPerm) {
    postorder(nb, bparent, post);
    std::vector<int> order2(nb);
    for (int k = 0; k < nb; k++) order2[k] = order[post[k]];
    order.swap(order2);
    for (int J = 0; J < nb; J++) pos[order[J]] = J;
    etree_graph(nb, g.ptr, g.adj, pos, order, bparent);
  }
  // children lists
  std::vector<int> chead(nb, -1), cnext(nb, -1), nchild(nb, 0);
  for (int J = nb - 1; J >= 0; J--)
    if (bparent[J] != -1) { cnext[J] = chead[bparent[J]]; chead[bparent[J]] = J; nchild[bparent[J]]++; }
  // block column structures by merging children (rows > J, in new numbering, sorted)
  std::vector<long long> sptr(nb + 1, 0);
  std::vector<int> spool;
  spool.reserve((size_t)g.adj.size() * 2);
  std::vector<int> stamp(nb, -1), tmp;
  for (int J = 0; J < nb; J++) {
    tmp.clear();
    stamp[J] = J;
    int v = order[J];
    for (int p = g.ptr[v]; p < g.ptr[v + 1]; p++) {
      int I = pos[g.adj[p]];
      if (I > J && stamp[I] != J) { stamp[I] = J; tmp.push_back(I); }
    }
    for (int C = chead[J]; C != -1; C = cnext[C])
      for (long long q = sptr[C]; q < sptr[C + 1]; q++) {
        int I = spool[q];
        if (I > J && stamp[I] != J) { stamp[I] = J; tmp.push_back(I); }
      }
    std::sort(tmp.begin(), tmp.end());
    spool.insert(spool.end(), tmp.begin(), tmp.end());
    sptr[J + 1] = (long long)spool.size();
  }
  // scalar positions of blocks in pivot order
  std::vector<int> bsize(nb), bpos(nb + 1, 0);
  for (int J = 0; J < nb; J++) { bsize[J] = g.w[order[J]]; bpos[J + 1] = bpos[J] + bsize[J]; }
  // weighted structure sizes W(J) = b_J + sum of sizes below
  std::vector<long long> Wt(nb);
  for (int J = 0; J < nb; J++) {
    long long w = bsize[J];
    for (long long q = sptr[J]; q < sptr[J + 1]; q++) w += bsize[spool[q]];
    Wt[J] = w;
  }
  // scalar permutation, etree and column counts
  S.perm.resize(n);
  S.iperm.resize(n);
  S.parent.assign(n, -1);
  S.colcount.assign(n, 0);
  S.nnzL = 0;
  S.flops = 0;
  for (int J = 0; J < nb; J++) {
    int ob = order[J];
    for (int t = 0; t < bsize[J]; t++) {
      int k = bpos[J] + t;
      S.perm[k] = bfirst[ob] + t;
      S.parent[k] = t + 1 < bsize[J] ? k + 1 : (bparent[J] >= 0 ? bpos[bparent[J]] : -1);
      S.colcount[k] = (int)(Wt[J] - t);
      S.nnzL += S.colcount[k];
      S.flops += (double)S.colcount[k] * S.colcount[k];
    }
  }
  for (int k = 0; k < n; k++) S.iperm[S.perm[k]] = k;
  S.nblocks = nb;
  S.block_size = bsize;
  // ---- 4. supernodes: fundamental, then relaxed amalgamation of contiguous last children
  std::vector<int> sfirst;  // first block of each supernode
  for (int J = 0; J < nb; J++) {
    bool cont = J > 0 && bparent[J - 1] == J && nchild[J] == 1 && (sptr[J] - sptr[J - 1]) == (sptr[J + 1] - sptr[J]) + 1;
    if (!cont) sfirst.push_back(J);
  }
  int ns0 = (int)sfirst.size();
  sfirst.push_back(nb);
  std::vector<int> sn_of_block(nb);
  for (int s = 0; s < ns0; s++) for (int J = sfirst[s]; J < sfirst[s + 1]; J++) sn_of_block[J] = s;
  // per supernode: scalar cols k, row count m (scalar, exact union = own + struct of last block), exact nnz
  struct SN { int b0, b1; long long k, m, exact; int parent; bool alive; };
  std::vector<SN> sn(ns0);
  for (int s = 0; s < ns0; s++) {
    int b0 = sfirst[s], b1 = sfirst[s + 1];
    long long k = bpos[b1] - bpos[b0];
    long long below = Wt[b1 - 1] - bsize[b1 - 1];
    long long ex = 0;
    for (int J = b0; J < b1; J++) for (int t = 0; t < bsize[J]; t++) ex += Wt[J] - t;
    int pb = bparent[b1 - 1];
    sn[s] = SN{b0, b1, k, k + below, ex, pb >= 0 ? sn_of_block[pb] : -1, true};
  }
  // merge s into parent p when s ends right before p starts
  std::vector<int> redirect(ns0);
  std::iota(redirect.begin(), redirect.end(), 0);
  auto find = [&](int s) { while (redirect[s] != s) s = redirect[s] = redirect[redirect[s]]; return s; };
  for (int s = 0; s < ns0; s++) {
    if (sn[s].parent < 0) continue;
    int p = find(sn[s].parent);
    if (sn[s].b1 != sn[p].b0) continue;
    long long k = sn[s].k + sn[p].k;
    long long m = sn[s].k + sn[p].m;
    long long panel = k * m - k * (k - 1) / 2;
    long long exact = sn[s].exact + sn[p].exact;
    double zf = panel > 0 ? double(panel - exact) / double(panel) : 0.0;
    bool ok = k <= opt.relax_small || (k <= opt.relax_max && zf <= opt.relax_zeros);
    if (!ok) continue;
    sn[p].b0 = sn[s].b0;
    sn[p].k = k;
    sn[p].m = m;
    sn[p].exact = exact;
    sn[s].alive = false;
    redirect[s] = p;
  }
  std::vector<int> newid(ns0, -1);
  int ns = 0;
  for (int s = 0; s < ns0; s++) if (sn[s].alive) newid[s] = ns++;
  S.ns = ns;
  S.sn_first.assign(ns + 1, 0);
  S.sn_parent.assign(ns, -1);
  S.sn_rowptr.assign(ns + 1, 0);
  S.col_to_sn.assign(n, -1);
  for (int s = 0; s < ns0; s++) {
    if (!sn[s].alive) continue;
    int id = newid[s];
    S.sn_first[id] = bpos[sn[s].b0];
    S.sn_parent[id] = sn[s].parent >= 0 ? newid[find(sn[s].parent)] : -1;
  }
  S.sn_first[ns] = n;
  for (int s = 0; s < ns; s++)
    for (int c = S.sn_first[s]; c < S.sn_first[s + 1]; c++) S.col_to_sn[c] = s;
  // row lists: own scalar columns, then the scalars of the blocks below the last block (superset of every column's)
  for (int s0 = 0; s0 < ns0; s0++) {
    if (!sn[s0].alive) continue;
    int s = newid[s0];
    // union of the below-structures of all blocks in the supernode (amalgamated supernodes may need more than the
    // last block's structure only when a merged child's structure is not nested; it always is for etree children)
    int blast = sn[s0].b1 - 1;
    for (int c = S.sn_first[s]; c < S.sn_first[s + 1]; c++) S.sn_rows.push_back(c);
    for (long long q = sptr[blast]; q < sptr[blast + 1]; q++) {
      int I = spool[q];
      for (int t = 0; t < bsize[I]; t++) S.sn_rows.push_back(bpos[I] + t);
    }
    S.sn_rowptr[s + 1] = (long long)S.sn_rows.size();
  }
  // verify nesting: every block's below-structure within the supernode's row list (cheap check on a sample is not
  // enough for correctness, so check all columns' structures fully)
  for (int s = 0; s < ns; s++) {
    const int* rb = &S.sn_rows[S.sn_rowptr[s]];
    const int* re = &S.sn_rows[0] + S.sn_rowptr[s + 1];
    int b0 = -1;
    // blocks of this supernode: those whose first scalar lies in [first, first+k)
    (void)b0;
    for (int c = S.sn_first[s]; c < S.sn_first[s + 1];) {
      // find block at scalar c
      int J = (int)(std::upper_bound(bpos.begin(), bpos.end(), c) - bpos.begin()) - 1;
      for (long long q = sptr[J]; q < sptr[J + 1]; q++) {
        int r = bpos[spool[q]];
        if (!std::binary_search(rb, re, r)) { S.error = "supernode structure not nested"; return -3; }
      }
      c = bpos[J + 1];
    }
  }
  std::vector<int>().swap(spool);
  // levels
  S.sn_level.assign(ns, 0);
  for (int s = 0; s < ns; s++)  // postorder: children before parents
    if (S.sn_parent[s] >= 0) S.sn_level[S.sn_parent[s]] = std::max(S.sn_level[S.sn_parent[s]], S.sn_level[s] + 1);
  S.max_level = 0;
  for (int s = 0; s < ns; s++) S.max_level = std::max(S.max_level, S.sn_level[s]);
  // ---- 5. storage offsets and maps
  // L panels: static. Update blocks: a pool reused across levels; U(s) lives from level(s) (allocated with its
  // level, before that level's extend-add) until the extend-add of level(parent(s)) has read it.
  S.Loff.assign(ns + 1, 0);
  S.Uoff.assign(ns + 1, 0);
  S.Voff.assign(ns + 1, 0);
  S.Usum = 0;
  for (int s = 0; s < ns; s++) {
    long long k = S.sn_first[s + 1] - S.sn_first[s];
    long long m = S.sn_rowptr[s + 1] - S.sn_rowptr[s];
    S.Loff[s + 1] = S.Loff[s] + k * m;
    S.Voff[s + 1] = S.Voff[s] + (m - k);
    S.Usum += (m - k) * (m - k);
  }
  S.Lsize = S.Loff[ns];
  S.Vsize = S.Voff[ns];
  {
    // first-fit over a free list of [offset, size) holes, level by level
    std::vector<std::vector<int>> bylev(S.max_level + 1);
    for (int s = 0; s < ns; s++) bylev[S.sn_level[s]].push_back(s);
    std::vector<std::pair<long long, long long>> holes;  // sorted by offset, coalesced
    long long top = 0;
    auto alloc = [&](long long sz) -> long long {
      if (sz <= 0) return 0;
      for (size_t h = 0; h < holes.size(); h++)
        if (holes[h].second >= sz) {
          long long off = holes[h].first;
          holes[h].first += sz;
          holes[h].second -= sz;
          if (holes[h].second == 0) holes.erase(holes.begin() + h);
          return off;
        }
      if (!holes.empty() && holes.back().first + holes.back().second == top) {  // grow the last hole
        long long off = holes.back().first;
        top = off + sz;
        holes.pop_back();
        return off;
      }
      long long off = top;
      top += sz;
      return off;
    };
    auto release = [&](long long off, long long sz) {
      if (sz <= 0) return;
      auto it = std::lower_bound(holes.begin(), holes.end(), std::make_pair(off, (long long)0));
      it = holes.insert(it, {off, sz});
      size_t i = it - holes.begin();
      if (i + 1 < holes.size() && holes[i].first + holes[i].second == holes[i + 1].first) {
        holes[i].second += holes[i + 1].second;
        holes.erase(holes.begin() + i + 1);
      }
      if (i > 0 && holes[i - 1].first + holes[i - 1].second == holes[i].first) {
        holes[i - 1].second += holes[i].second;
        holes.erase(holes.begin() + i);
      }
    };
    std::vector<long long> usz(ns);
    for (int s = 0; s < ns; s++) {
      long long k = S.sn_first[s + 1] - S.sn_first[s];
      long long m = S.sn_rowptr[s + 1] - S.sn_rowptr[s];
      usz[s] = (m - k) * (m - k);
    }
    std::vector<std::vector<int>> kids(ns);
    for (int s = 0; s < ns; s++) if (S.sn_parent[s] >= 0) kids[S.sn_parent[s]].push_back(s);
    for (int l = 0; l <= S.max_level; l++) {
      for (int s : bylev[l]) S.Uoff[s] = alloc(usz[s]);
      // the children of this level's supernodes are read by this level's extend-add, then their space is free
      for (int s : bylev[l])
        for (int c : kids[s]) release(S.Uoff[c], usz[c]);
    }
    S.Usize = std::max<long long>(top, 1);
  }
  S.rel_off.assign(ns + 1, 0);
  for (int s = 0; s < ns; s++) {
    long long k = S.sn_first[s + 1] - S.sn_first[s];
    long long m = S.sn_rowptr[s + 1] - S.sn_rowptr[s];
    int p = S.sn_parent[s];
    S.rel_off[s + 1] = S.rel_off[s] + (p >= 0 ? (m - k) : 0);
    if (p < 0) {
      if (m != k) { S.error = "root supernode with off-diagonal rows"; return -3; }
      continue;
    }
    const int* pb = &S.sn_rows[S.sn_rowptr[p]];
    const int* pe = &S.sn_rows[0] + S.sn_rowptr[p + 1];
    for (long long q = S.sn_rowptr[s] + k; q < S.sn_rowptr[s + 1]; q++) {
      const int* f = std::lower_bound(pb, pe, S.sn_rows[q]);
      if (f == pe || *f != S.sn_rows[q]) { S.error = "child row missing from parent"; return -3; }
      S.rel.push_back((int)(f - pb));
    }
  }
  // A -> L scatter map (entries of the lower triangle of P A P^T)
  S.amap_src.clear();
  S.amap_dst.clear();
  for (int j = 0; j < n; j++) {
    int pj = S.iperm[j];
    int s = S.col_to_sn[pj];
    long long k0 = S.sn_first[s];
    long long m = S.sn_rowptr[s + 1] - S.sn_rowptr[s];
    const int* rb = &S.sn_rows[S.sn_rowptr[s]];
    const int* re = &S.sn_rows[0] + S.sn_rowptr[s + 1];
    for (int p = colPtr[j]; p < colPtr[j + 1]; p++) {
      int pi = S.iperm[rowIdx[p]];
      if (pi < pj) continue;
      const int* f = std::lower_bound(rb, re, pi);
      if (f == re || *f != pi) { S.error = "A entry outside the factor structure"; return -3; }
      S.amap_src.push_back(p);
      S.amap_dst.push_back(S.Loff[s] + (pj - k0) * m + (f - rb));
    }
  }
  return 0;
}

}  // namespace ifea
