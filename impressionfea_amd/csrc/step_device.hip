// Device-resident solve step of the 3D rigid-viscoplastic solver (ImpressionFEA Forging3DSolver hot path) on MI355X.
//
// Data layout in HBM (structure of arrays, FP64):
//   per node      x[3nn], frames ax[9nn] (rows ax0 ax1 ax2), nfix, contact, v/r/dv[3nn] in local frames
//   per element   tets[4ne], grad[12ne], vol, eps, sigma, rate, dev[6ne], and per assembly wm = vol*mu,
//                 wn = vol*coefN*blend, G[12ne], energy, power
//   per pattern   the reference's full symmetric CSC (colPtr, rowIdx, val, kvol), 3x3 node blocks
// Assembly is a two-pass, atomic-free gather: an element pass (strain rate, flow stress, G = B^T M eps) and a node
// pass where one wave owns the 3 CSC columns of a node and adds its incident elements' rotated 3x3 blocks in
// ascending element order, as the reference's serial scatter does (Forging3DSolver.cs:1146-1171).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>

#include "step_device.h"

namespace ifea {

constexpr int MAX_DEG = 192;    // two-ring neighbours per node
constexpr int MAX_PATCH = 64;   // one-ring patch size per node
constexpr int NODE_WAVES = 4;   // nodes per 256-thread block
// dynamic LDS of the per-node assembly kernels for a mesh whose largest neighbour list has md nodes (sized to the
// mesh, not to MAX_DEG, so that more workgroups share a CU)
inline size_t node_lds_bytes(int md) { return (size_t)NODE_WAVES * md * (9 * sizeof(double) + sizeof(int)); }
constexpr int TAN_WAVES = 2;    // nodes per workgroup of the tangent assembly
constexpr int NFIX_BATCH = 4;   // column entries per lane (of 64) whose rows' fixed-DOF counts the assembly gathers at once
// dynamic LDS of the tangent assembly: per wave 18 md + 26 me doubles and md + 9 me ints (md: largest neighbour list,
// me: most elements around a node)
inline size_t node_tan_lds_bytes(int md, int me) {
  return (size_t)TAN_WAVES * ((18 * (size_t)md + 26 * (size_t)me) * sizeof(double) + ((size_t)md + 9 * (size_t)me) * sizeof(int));
}

__device__ __forceinline__ double law_stress(const ifea_step_params& P, double strain, double rate) {
  return P.C * pow(P.eps0 + fmax(strain, 0.0), P.n) * pow(fmax(rate, 1e-12), P.m);
}

__device__ __forceinline__ int bsearch_lds(const int* a, int n, int key) {
  int lo = 0, hi = n - 1;
  while (lo <= hi) {
    int mid = (lo + hi) >> 1;
    int v = a[mid];
    if (v == key) return mid;
    if (v < key) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double wsum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return __shfl(v, 0, 64);
}

// ------------------------------------------------------------------------------- PrepareStep (:730-753)
__global__ void k_geom(int ne, const int* __restrict__ tets, const double* __restrict__ x, double* __restrict__ grad,
                       double* __restrict__ vol, int* __restrict__ bad) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int* t = tets + 4 * e;
  double p0x = x[3 * t[0]], p0y = x[3 * t[0] + 1], p0z = x[3 * t[0] + 2];
  double ax = x[3 * t[1]] - p0x, ay = x[3 * t[1] + 1] - p0y, az = x[3 * t[1] + 2] - p0z;
  double bx = x[3 * t[2]] - p0x, by = x[3 * t[2] + 1] - p0y, bz = x[3 * t[2] + 2] - p0z;
  double cx = x[3 * t[3]] - p0x, cy = x[3 * t[3] + 1] - p0y, cz = x[3 * t[3] + 2] - p0z;
  // b x c, c x a, a x b
  double bcx = by * cz - bz * cy, bcy = bz * cx - bx * cz, bcz = bx * cy - by * cx;
  double cax = cy * az - cz * ay, cay = cz * ax - cx * az, caz = cx * ay - cy * ax;
  double abx = ay * bz - az * by, aby = az * bx - ax * bz, abz = ax * by - ay * bx;
  double det = ax * bcx + ay * bcy + az * bcz;
  if (!(det > 0)) atomicMin(bad, e);
  double inv = 1.0 / det;
  double g1x = inv * bcx, g1y = inv * bcy, g1z = inv * bcz;
  double g2x = inv * cax, g2y = inv * cay, g2z = inv * caz;
  double g3x = inv * abx, g3y = inv * aby, g3z = inv * abz;
  double* g = grad + 12 * (long long)e;
  g[0] = -(g1x + g2x + g3x); g[1] = -(g1y + g2y + g3y); g[2] = -(g1z + g2z + g3z);
  g[3] = g1x; g[4] = g1y; g[5] = g1z;
  g[6] = g2x; g[7] = g2y; g[8] = g2z;
  g[9] = g3x; g[10] = g3y; g[11] = g3z;
  vol[e] = det / 6;
}

// BuildAnp (:761-823): nodal volumes, nodal gradient operator, penalty matrix
__global__ void k_vnode(int nn, const int* __restrict__ nts, const int* __restrict__ ntl, const double* __restrict__ vol,
                        double* __restrict__ vnode) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  double s = 0.0;
  for (int q = nts[n]; q < nts[n + 1]; q++) s += vol[ntl[q]] / 4;
  vnode[n] = s;
}

__global__ void k_gcoef(int nn, const int* __restrict__ tets, const int* __restrict__ nts, const int* __restrict__ ntl,
                        const int* __restrict__ ps, const int* __restrict__ pn, const double* __restrict__ vol,
                        const double* __restrict__ vnode, const double* __restrict__ grad, double* __restrict__ gcoef) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  const int p0 = ps[n], np = ps[n + 1] - p0;
  for (int t = 0; t < 3 * np; t++) gcoef[3 * (long long)p0 + t] = 0.0;
  for (int q = nts[n]; q < nts[n + 1]; q++) {
    int e = ntl[q];
    double w = vol[e] / 4 / vnode[n];
    for (int a = 0; a < 4; a++) {
      int m = tets[4 * e + a];
      int lo = 0, hi = np - 1, sl = 0;
      while (lo <= hi) { int mid = (lo + hi) >> 1; int v = pn[p0 + mid]; if (v == m) { sl = mid; break; } if (v < m) lo = mid + 1; else hi = mid - 1; }
      const double* g = grad + 12 * (long long)e + 3 * a;
      double* c = gcoef + 3 * ((long long)p0 + sl);
      c[0] += w * g[0]; c[1] += w * g[1]; c[2] += w * g[2];
    }
  }
}

__global__ __launch_bounds__(256) void k_kvol(int md, int nn, double kpen, const int* __restrict__ colPtr,
                                              const int* __restrict__ nbs, const int* __restrict__ nb,
                                              const int* __restrict__ ps, const int* __restrict__ pn,
                                              const double* __restrict__ ax, const double* __restrict__ vnode,
                                              const double* __restrict__ gcoef, double* __restrict__ kvol) {
  extern __shared__ double node_lds[];  // node_lds_bytes(md): per wave 9 md accumulators, then md neighbour ids
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int j = blockIdx.x * NODE_WAVES + wv;
  if (j >= nn) return;
  const int deg = nbs[j + 1] - nbs[j];
  double* acc = node_lds + (long long)wv * 9 * md;
  int* nbr = (int*)(node_lds + (long long)NODE_WAVES * 9 * md) + wv * md;
  for (int q = lane; q < deg; q += 64) nbr[q] = nb[nbs[j] + q];
  for (int q = lane; q < 9 * deg; q += 64) acc[q] = 0.0;
  wave_sync_lds();
  const double* Rj = ax + 9 * (long long)j;
  const int d3 = 3 * deg;
  for (int pp = ps[j]; pp < ps[j + 1]; pp++) {
    const int n = pn[pp];
    const double kv = kpen * vnode[n];
    const int p0 = ps[n], np = ps[n + 1] - p0;
    int sj = 0;
    {
      int lo = 0, hi = np - 1;
      while (lo <= hi) { int mid = (lo + hi) >> 1; int v = pn[p0 + mid]; if (v == j) { sj = mid; break; } if (v < j) lo = mid + 1; else hi = mid - 1; }
    }
    const double* gj = gcoef + 3 * ((long long)p0 + sj);
    double cj[3];
    for (int c = 0; c < 3; c++) cj[c] = gj[0] * Rj[3 * c] + gj[1] * Rj[3 * c + 1] + gj[2] * Rj[3 * c + 2];
    if (lane < np) {
      const int i = pn[p0 + lane];
      const double* gi = gcoef + 3 * ((long long)p0 + lane);
      const double* Ri = ax + 9 * (long long)i;
      double ci[3];
      for (int r = 0; r < 3; r++) ci[r] = gi[0] * Ri[3 * r] + gi[1] * Ri[3 * r + 1] + gi[2] * Ri[3 * r + 2];
      const int pos = bsearch_lds(nbr, deg, i);
      if (pos >= 0) {
        for (int c = 0; c < 3; c++)
          for (int r = 0; r < 3; r++) acc[c * d3 + 3 * pos + r] += kv * (ci[r] * cj[c]);
      }
    }
    wave_sync_lds();
  }
  for (int c = 0; c < 3; c++) {
    double* dst = kvol + colPtr[3 * j + c];
    for (int q = lane; q < d3; q += 64) dst[q] = acc[c * d3 + q];
  }
}

// BuildFriction (:829-850): faces of each node in the reference's face order
__global__ void k_fric(int nn, const int* __restrict__ nfs, const int* __restrict__ nfl, const int* __restrict__ faces,
                       const int* __restrict__ owner, const double* __restrict__ x, const double* __restrict__ sig,
                       double* __restrict__ afric, double* __restrict__ kfric) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  double a = 0.0, ks = 0.0;
  int cnt = 0;
  for (int q = nfs[n]; q < nfs[n + 1]; q++) {
    int f = nfl[q];
    int ia = faces[3 * f], ib = faces[3 * f + 1], ic = faces[3 * f + 2];
    double ux = x[3 * ib] - x[3 * ia], uy = x[3 * ib + 1] - x[3 * ia + 1], uz = x[3 * ib + 2] - x[3 * ia + 2];
    double wx = x[3 * ic] - x[3 * ia], wy = x[3 * ic + 1] - x[3 * ia + 1], wz = x[3 * ic + 2] - x[3 * ia + 2];
    double cx = uy * wz - uz * wy, cy = uz * wx - ux * wz, cz = ux * wy - uy * wx;
    double area = 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
    a += area / 3;
    ks += sig[owner[f]] / sqrt(3.0);
    cnt++;
  }
  afric[n] = a;
  kfric[n] = cnt > 0 ? ks / cnt : 0.0;
}

// ------------------------------------------------------------------------------- PrepareStep frames (:669-727)
// Analytic rigid dies as height fields (the host's DieSpec and the oracle's Die::height): kind 0 a flat plane at
// z = p[0], kind 1 a top die with a hexagonal cavity, kind 2 one half of a connecting-rod closed die. The frames are
// evaluated with the same IEEE operations in the same order as on the host (no FMA contraction in these functions),
// so they agree bit for bit.
__device__ double die_smooth(double u) {
#pragma clang fp contract(off)
  u = fmin(1.0, fmax(0.0, u));
  return u * u * (3 - 2 * u);
}

__device__ double die_height(const DieDev& D, double x, double y) {
#pragma clang fp contract(off)
  const double* p = D.p;
  if (D.kind == 0) return p[0];
  if (D.kind == 2) {
    const double xb = -42.0, rb = 20.0, xs = 49.5, rs = 12.5, g = 5.0;
    const double t = p[0], wall = p[6], land = p[7], w = p[5], fl = 0.25 * w;
    const double db = sqrt((x - xb) * (x - xb) + y * y) - rb, ds = sqrt((x - xs) * (x - xs) + y * y) - rs;
    const double dsh = fmax(fmax(xb - x, x - xs), fabs(y) - 0.5 * w);
    const double shank = p[4] + (p[3] - p[4]) * die_smooth((fabs(y) - (0.5 * w - fl) + 0.5 * wall) / wall);
    const double depth = fmax(fmax(p[1] * die_smooth((-db + 0.5 * wall) / wall), p[2] * die_smooth((-ds + 0.5 * wall) / wall)),
                              shank * die_smooth((-dsh + 0.5 * wall) / wall));
    const double h = 0.5 * t + depth + g * die_smooth((fmin(fmin(db, ds), dsh) - land) / wall);
    return D.side == 1 ? h : -h;
  }
  const double s3 = sqrt(3.0) / 2;
  const double d = fmax(fabs(x), fmax(fabs(0.5 * x + s3 * y), fabs(-0.5 * x + s3 * y))) - 0.5 * p[2];
  double h = p[0] + p[1] * die_smooth((-d + 0.5 * p[3]) / p[3]);
  h += p[5] * die_smooth((d - p[4]) / p[3]);
  return h;
}

// normal of the die surface below (x, y), pointing out of the die: central differences of the height
// (HeightFieldDie.Project's normal, HeightFieldDie.cs:157-191, for an analytic surface)
__device__ void die_normal(const DieDev& D, double x, double y, double n[3]) {
#pragma clang fp contract(off)
  const double dd = 1e-5;
  const double hx = (die_height(D, x + dd, y) - die_height(D, x - dd, y)) / (2 * dd);
  const double hy = (die_height(D, x, y + dd) - die_height(D, x, y - dd)) / (2 * dd);
  double u0 = -hx, u1 = -hy, u2 = 1.0;
  const double l = sqrt(u0 * u0 + u1 * u1 + u2 * u2);
  if (l > 0) { u0 = u0 / l; u1 = u1 / l; u2 = u2 / l; }
  n[0] = D.side == 0 ? u0 : -u0;
  n[1] = D.side == 0 ? u1 : -u1;
  n[2] = D.side == 0 ? u2 : -u2;
}

__device__ __forceinline__ double dot3(const double a[3], const double b[3]) {
#pragma clang fp contract(off)
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

__device__ __forceinline__ void cross3(const double a[3], const double b[3], double c[3]) {
#pragma clang fp contract(off)
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

__device__ __forceinline__ void normalize3(double a[3]) {
#pragma clang fp contract(off)
  const double l = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  if (l > 0) { a[0] = a[0] / l; a[1] = a[1] / l; a[2] = a[2] / l; }
}

// Per node: constraint directions (symmetry planes x = 0 and y = 0, the die normal of a contact node), Gram-Schmidt
// with the 0.2 drop threshold, completion to an orthonormal frame; the first nfix local DOFs are fixed to the
// normal die velocity (:669-727).
__global__ void k_frames(int nn, DieDev bottom, DieDev top, double ram, const double* __restrict__ x,
                         const unsigned char* __restrict__ on_x0, const unsigned char* __restrict__ on_y0,
                         const int* __restrict__ contact, double* __restrict__ ax, int* __restrict__ nfix,
                         int* __restrict__ die_axis, double* __restrict__ fixed_val) {
#pragma clang fp contract(off)
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  double dirs[3][3], vals[3], e[3][3], ev[3];
  int k = 0, die_idx = -1;
#pragma unroll
  for (int q = 0; q < 3; q++) {
    vals[q] = 0.0;
    ev[q] = 0.0;
#pragma unroll
    for (int c = 0; c < 3; c++) { dirs[q][c] = 0.0; e[q][c] = 0.0; }
  }
  if (on_x0[n]) { dirs[k][0] = 1.0; k++; }
  if (on_y0[n]) {
    if (k == 0) dirs[0][1] = 1.0; else dirs[1][1] = 1.0;
    k++;
  }
  const int ct = contact[n];
  if (ct != 0) {
    double nrm[3];
    die_normal(ct == 2 ? top : bottom, x[3 * n], x[3 * n + 1], nrm);
    // DieVelocity . nrm: the top die moves with (0, 0, -ram), the bottom die rests
    const double v = ct == 2 ? 0.0 * nrm[0] + 0.0 * nrm[1] + (-ram) * nrm[2] : 0.0 * nrm[0] + 0.0 * nrm[1] + 0.0 * nrm[2];
#pragma unroll
    for (int q = 0; q < 3; q++)
      if (q == k) { dirs[q][0] = nrm[0]; dirs[q][1] = nrm[1]; dirs[q][2] = nrm[2]; vals[q] = v; }
    die_idx = k;
    k++;
  }
  int m = 0, da = -1;
#pragma unroll
  for (int q = 0; q < 3; q++) {
    if (q >= k) continue;
    double d[3] = {dirs[q][0], dirs[q][1], dirs[q][2]};
    double vv = vals[q];
#pragma unroll
    for (int s2 = 0; s2 < 3; s2++) {
      if (s2 >= m) continue;
      const double proj = dot3(d, e[s2]);
      d[0] = d[0] - proj * e[s2][0];
      d[1] = d[1] - proj * e[s2][1];
      d[2] = d[2] - proj * e[s2][2];
      vv -= proj * ev[s2];
    }
    const double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (len < 0.2) continue;
    const double il = 1 / len;
#pragma unroll
    for (int t = 0; t < 3; t++)
      if (t == m) { e[t][0] = il * d[0]; e[t][1] = il * d[1]; e[t][2] = il * d[2]; ev[t] = vv / len; }
    if (q == die_idx) da = m;
    m++;
  }
  if (m == 0) {
#pragma unroll
    for (int q = 0; q < 3; q++)
#pragma unroll
      for (int c = 0; c < 3; c++) e[q][c] = q == c ? 1.0 : 0.0;
  } else if (m == 1) {
    const double helper[3] = {fabs(e[0][0]) < 0.8 ? 1.0 : 0.0, fabs(e[0][0]) < 0.8 ? 0.0 : 1.0, 0.0};
    cross3(e[0], helper, e[1]);
    normalize3(e[1]);
    cross3(e[0], e[1], e[2]);
  } else if (m == 2) {
    cross3(e[0], e[1], e[2]);
    normalize3(e[2]);
  }
  double* R = ax + 9 * (long long)n;
#pragma unroll
  for (int q = 0; q < 3; q++)
#pragma unroll
    for (int c = 0; c < 3; c++) R[3 * q + c] = e[q][c];
  nfix[n] = m;
  die_axis[n] = da;
#pragma unroll
  for (int q = 0; q < 3; q++) fixed_val[3 * n + q] = q < m ? ev[q] : 0.0;
}

// LoadVelocityGuess (:854-869)
__global__ void k_vguess(int nn, const double* __restrict__ ax, const int* __restrict__ nfix,
                         const double* __restrict__ fixed_val, const double* __restrict__ vg, double* __restrict__ v) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  const double* R = ax + 9 * (long long)n;
  for (int q = 0; q < 3; q++) {
    double l = vg[3 * n] * R[3 * q] + vg[3 * n + 1] * R[3 * q + 1] + vg[3 * n + 2] * R[3 * q + 2];
    v[3 * n + q] = q < nfix[n] ? fixed_val[3 * n + q] : l;
  }
}

// ------------------------------------------------------------------------------- Kernel (:1312-1400), element pass
__global__ __launch_bounds__(256) void k_elem(int ne, ifea_step_params P, int newton, const int* __restrict__ tets,
                                              const double* __restrict__ ax, const double* __restrict__ v,
                                              const double* __restrict__ grad, const double* __restrict__ vol,
                                              const double* __restrict__ eps, double* __restrict__ wm_out,
                                              double* __restrict__ wn_out, double* __restrict__ g_out,
                                              double* __restrict__ energy_out, double* __restrict__ power_out,
                                              double* __restrict__ el_sig, double* __restrict__ el_rate,
                                              double* __restrict__ el_dev) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  double ve[12], g[12];
  const double* ge = grad + 12 * (long long)e;
#pragma unroll
  for (int k = 0; k < 12; k++) g[k] = ge[k];
#pragma unroll
  for (int a = 0; a < 4; a++) {
    int n = tets[4 * e + a];
    const double* R = ax + 9 * (long long)n;
    double l0 = v[3 * n], l1 = v[3 * n + 1], l2 = v[3 * n + 2];
    // ToGlobal (:827): l0*ax0 + l1*ax1 + l2*ax2
    ve[3 * a] = l0 * R[0] + l1 * R[3] + l2 * R[6];
    ve[3 * a + 1] = l0 * R[1] + l1 * R[4] + l2 * R[7];
    ve[3 * a + 2] = l0 * R[2] + l1 * R[5] + l2 * R[8];
  }
  // eps = B ve (engineering shear)
  double e0 = 0, e1 = 0, e2 = 0, e3 = 0, e4 = 0, e5 = 0;
#pragma unroll
  for (int a = 0; a < 4; a++) {
    double gx = g[3 * a], gy = g[3 * a + 1], gz = g[3 * a + 2];
    double vx = ve[3 * a], vy = ve[3 * a + 1], vz = ve[3 * a + 2];
    e0 += gx * vx;
    e1 += gy * vy;
    e2 += gz * vz;
    e3 += gy * vx + gx * vy;
    e4 += gz * vy + gy * vz;
    e5 += gz * vx + gx * vz;
  }
  double ev3 = (e0 + e1 + e2) / 3;
  double d0 = e0 - ev3, d1 = e1 - ev3, d2 = e2 - ev3;
  double rate = sqrt(fmax(2.0 / 3.0 * (d0 * d0 + d1 * d1 + d2 * d2) + (e3 * e3 + e4 * e4 + e5 * e5) / 3.0, 0.0));
  double me0 = 2.0 / 3.0 * d0, me1 = 2.0 / 3.0 * d1, me2 = 2.0 / 3.0 * d2, me3 = e3 / 3, me4 = e4 / 3, me5 = e5 / 3;
  double strain = eps[e];
  double sigCut = law_stress(P, strain, P.rate_cut);
  double mu0 = sigCut / P.rate_cut;
  double mu, sig, wd, coefN;
  if (rate > P.rate_cut) {
    sig = law_stress(P, strain, rate);
    mu = sig / rate;
    wd = 0.5 * sigCut * P.rate_cut + (sig * rate - sigCut * P.rate_cut) / (P.m + 1);
    coefN = (P.m - 1) * mu / (rate * rate);
  } else {
    mu = mu0;
    sig = mu0 * rate;
    wd = 0.5 * mu0 * rate * rate;
    coefN = 0;
  }
  double ve_ = vol[e];
  energy_out[e] = ve_ * wd;
  power_out[e] = ve_ * sig * rate;
  el_sig[e] = sig;
  el_rate[e] = rate;
  double* dv = el_dev + 6 * (long long)e;
  dv[0] = mu * me0; dv[1] = mu * me1; dv[2] = mu * me2; dv[3] = mu * me3; dv[4] = mu * me4; dv[5] = mu * me5;
  wm_out[e] = ve_ * mu;
  wn_out[e] = newton ? ve_ * coefN * P.newton_blend : 0.0;
  double* G = g_out + 12 * (long long)e;
#pragma unroll
  for (int a = 0; a < 4; a++) {
    double gx = g[3 * a], gy = g[3 * a + 1], gz = g[3 * a + 2];
    G[3 * a] = gx * me0 + gy * me3 + gz * me5;
    G[3 * a + 1] = gy * me1 + gx * me3 + gz * me4;
    G[3 * a + 2] = gz * me2 + gy * me4 + gx * me5;
  }
}

// ------------------------------------------------------------------------------- Assemble (:1115-1217), node pass
// One wave per node j: its 3 CSC columns (and its 3 residual rows by symmetry). mode 0: raw residual; mode 1:
// BC identity rows/cols, r[fixed] = 0 and r = -r (SolveVelocity :892-917); mode 2: energy only.
// Friction of a contact node on its free axes (Forging3DSolver.cs:1175-1214): energy, power and residual terms added
// to en, pw, rj; the slip quantities are returned for the tangent.
struct NodeFriction {
  bool on;
  double c, s[3], s2, sl, z, gs;
};
__device__ __forceinline__ NodeFriction node_friction(const ifea_step_params& P, int j, int nf, int ct,
                                                      const double* __restrict__ Rj, const double* __restrict__ v,
                                                      const double* __restrict__ afric,
                                                      const double* __restrict__ kfric, double& en, double& pw,
                                                      double rj[3]) {
  NodeFriction f{};
  f.on = false;
  if (P.friction > 0 && ct != 0 && nf < 3) {
    const double c = P.friction * kfric[j] * afric[j] * (2 / M_PI);
    if (c > 0) {
      // die velocity in local axes: top die moves with (0, 0, -ram)
      double vd[3];
      for (int qq = 0; qq < 3; qq++) vd[qq] = ct == 2 ? -P.ram * Rj[3 * qq + 2] : 0.0;
      double s2 = 0;
      for (int qq = 0; qq < 3; qq++) f.s[qq] = 0.0;
      for (int qq = nf; qq < 3; qq++) { f.s[qq] = v[3 * j + qq] - vd[qq]; s2 += f.s[qq] * f.s[qq]; }
      const double sl = sqrt(s2), z = sl / P.u0, at = atan(z);
      en += c * (sl * at - 0.5 * P.u0 * log(1 + z * z));
      pw += c * at * sl;
      const double gs = sl > 1e-12 * P.u0 ? at / sl : 1 / P.u0;
      for (int qq = nf; qq < 3; qq++) rj[qq] += c * gs * f.s[qq];
      f.on = true;
      f.c = c;
      f.s2 = s2;
      f.sl = sl;
      f.z = z;
      f.gs = gs;
    }
  }
  return f;
}

// Tangent assembly (with the residual, energy and power of the same call): one wave per node j, TAN_WAVES nodes per
// workgroup, node_tan_lds_bytes(md, me) of dynamic LDS (per wave: the three CSC columns of j as 3x3 blocks, the frames
// of j's neighbours, and the data of the elements around j). Everything an element contributes is staged first with
// all loads in flight (one element per lane), so the element loop, which accumulates in ascending element order,
// reads LDS only.
__global__ __launch_bounds__(64 * TAN_WAVES) void k_node(int md, int me, int nn, ifea_step_params P, int matrix,
                                                         int newton, int mode, const int* __restrict__ tets,
                                                         const int* __restrict__ nts, const int* __restrict__ ntl,
                                                         const int* __restrict__ colPtr,
                                                         const int* __restrict__ rowIdx, const int* __restrict__ nbs,
                                                         const int* __restrict__ nb, const double* __restrict__ ax,
                                                         const int* __restrict__ nfix, const int* __restrict__ contact,
                                                         const double* __restrict__ afric,
                                                         const double* __restrict__ kfric,
                                                         const double* __restrict__ kvol, const double* __restrict__ v,
                                                         const double* __restrict__ grad,
                                                         const double* __restrict__ wm, const double* __restrict__ wn,
                                                         const double* __restrict__ G, double* __restrict__ r,
                                                         double* __restrict__ val, double* __restrict__ n_energy,
                                                         double* __restrict__ n_power,
                                                         unsigned long long* __restrict__ maxdiag) {
  extern __shared__ double node_lds[];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int j = blockIdx.x * TAN_WAVES + wv;
  if (j >= nn) return;
  const int deg = nbs[j + 1] - nbs[j];
  const int d3 = 3 * deg;
  const int q0 = nts[j], nej = nts[j + 1] - q0;  // elements around j, nej <= me
  double* acc = node_lds + (long long)wv * (18 * md + 26 * me);  // 9 md: the node's columns, [c][3 pos + r]
  double* axn = acc + 9 * md;                                     // 9 md: frames of the neighbours
  double* eg = axn + 9 * md;                                      // 12 me: element gradients
  double* eG = eg + 12 * me;                                      // 12 me: element G vectors
  double* ew = eG + 12 * me;                                      // me: wm
  double* ewn = ew + me;                                          // me: wn
  int* nbr = (int*)(node_lds + (long long)TAN_WAVES * (18 * md + 26 * me)) + wv * (md + 9 * me);
  int* et = nbr + md;                                             // 4 me: neighbour positions of the element's nodes
  int* eb = et + 4 * me;                                          // me: local index of j in the element
  // The staging loops below issue every load of a batch before using any: a node's work is a few short chains of
  // dependent loads (index, then gather), and one wave per node leaves the latency exposed unless the chains overlap.
  const int* nbj = nb + nbs[j];
  for (int q = lane; q < deg; q += 64) nbr[q] = nbj[q];
  // the node's three columns of K_vol into acc, and r_vol = K_vol v (row j = column j, symmetric), each lane summing
  // its entries in ascending order
  double rj[3];
  {
    constexpr int KB = 4;  // entries of a column per lane per batch
    double s[3] = {0.0, 0.0, 0.0};
    for (int base = 0; base < d3; base += 64 * KB) {
      double kv[3][KB], vv[3][KB];
      int ri[3][KB];
#pragma unroll
      for (int c = 0; c < 3; c++)
#pragma unroll
        for (int k = 0; k < KB; k++) {
          const int q = base + lane + 64 * k;
          const int p = colPtr[3 * j + c] + min(q, d3 - 1);
          kv[c][k] = kvol[p];
          ri[c][k] = rowIdx[p];
        }
#pragma unroll
      for (int c = 0; c < 3; c++)
#pragma unroll
        for (int k = 0; k < KB; k++) vv[c][k] = v[ri[c][k]];
#pragma unroll
      for (int c = 0; c < 3; c++)
#pragma unroll
        for (int k = 0; k < KB; k++) {
          const int q = base + lane + 64 * k;
          if (q < d3) {
            acc[c * d3 + q] = kv[c][k];
            s[c] += kv[c][k] * vv[c][k];
          }
        }
    }
#pragma unroll
    for (int c = 0; c < 3; c++) rj[c] = wsum(s[c]);
  }
  // the frames of the neighbours
  {
    constexpr int KB = 8;
    for (int base = 0; base < 9 * deg; base += 64 * KB) {
      int nbk[KB];
      double av[KB];
#pragma unroll
      for (int k = 0; k < KB; k++) nbk[k] = nbj[min(base + lane + 64 * k, 9 * deg - 1) / 9];
#pragma unroll
      for (int k = 0; k < KB; k++) {
        const int q = min(base + lane + 64 * k, 9 * deg - 1);
        av[k] = ax[9 * (long long)nbk[k] + q % 9];
      }
#pragma unroll
      for (int k = 0; k < KB; k++) {
        const int q = base + lane + 64 * k;
        if (q < 9 * deg) axn[q] = av[k];
      }
    }
  }
  double en = 0.5 * v[3 * j] * rj[0] + 0.5 * v[3 * j + 1] * rj[1] + 0.5 * v[3 * j + 2] * rj[2];
  double pw = 2.0 * en;
  const double* Rj = ax + 9 * (long long)j;
  double Rjv[9];
#pragma unroll
  for (int u = 0; u < 9; u++) Rjv[u] = Rj[u];
  wave_sync_lds();  // nbr complete before the searches below
  // stage the elements, one per lane, and add their residual terms ToLocal(j, wm G_b) (wave sum in a fixed order)
  double c3[3] = {0.0, 0.0, 0.0};
  for (int q = lane; q < nej; q += 64) {
    const int e = ntl[q0 + q];
    int tt[4];
#pragma unroll
    for (int a = 0; a < 4; a++) tt[a] = tets[4 * e + a];
    const int b = tt[0] == j ? 0 : (tt[1] == j ? 1 : (tt[2] == j ? 2 : 3));
    const double w = wm[e];
    eb[q] = b;
    ew[q] = w;
    ewn[q] = wn[e];
    const double* ge = grad + 12 * (long long)e;
    const double* Ge = G + 12 * (long long)e;
    double gv[12], Gv[12];
#pragma unroll
    for (int u = 0; u < 12; u++) {
      gv[u] = ge[u];
      Gv[u] = Ge[u];
    }
#pragma unroll
    for (int u = 0; u < 12; u++) {
      eg[12 * q + u] = gv[u];
      eG[12 * q + u] = Gv[u];
    }
#pragma unroll
    for (int a = 0; a < 4; a++) et[4 * q + a] = bsearch_lds(nbr, deg, tt[a]);
    const double reb0 = w * Gv[3 * b], reb1 = w * Gv[3 * b + 1], reb2 = w * Gv[3 * b + 2];
#pragma unroll
    for (int c = 0; c < 3; c++) c3[c] += reb0 * Rjv[3 * c] + reb1 * Rjv[3 * c + 1] + reb2 * Rjv[3 * c + 2];
  }
#pragma unroll
  for (int c = 0; c < 3; c++) rj[c] += wsum(c3[c]);
  wave_sync_lds();
  // elements around j in ascending order, 36 lanes per element for the 4 x 3 x 3 block entries. Lane (a, c, d) =
  // 9 a + 3 c + d computes the element matrix entry Ke_ab[c][d] once; the rotated block
  // blk[r][c'] = sum_c Ri[r][c] (sum_d Ke[c][d] Rj[c'][d]) then takes two three-term products, the inner one from
  // the row (a, c, *) and the outer one from the column (a, *, c') of the lanes' values (lane shuffles)
  const int la = lane / 9, lr = (lane / 3) % 3, lc = lane % 3;
  const int lac = min(la, 3);  // lanes past the 36 compute a copy of block 3 and store nothing
  const double rl0 = lc == 0 ? Rjv[0] : (lc == 1 ? Rjv[3] : Rjv[6]);
  const double rl1 = lc == 0 ? Rjv[1] : (lc == 1 ? Rjv[4] : Rjv[7]);
  const double rl2 = lc == 0 ? Rjv[2] : (lc == 1 ? Rjv[5] : Rjv[8]);
  const int src_row = 9 * lac + 3 * lr, src_col = 9 * lac + lc;
  for (int q = 0; q < nej; q++) {
    const int pos = lane < 36 ? et[4 * q + la] : -1;
    const int b = eb[q];
    const double w = ew[q], wnn = ewn[q];
    const double* ga = eg + 12 * q + 3 * lac;
    const double* gb = eg + 12 * q + 3 * b;
    const double* Gq = eG + 12 * q;
    const double gab = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
    // Ke_ab[c][d] = wm * (B^T M B)_{3a+c, 3b+d} + wn * G_{3a+c} G_{3b+d}, with c = lr, d = lc
    const double ga_c = ga[lr], gb_d = gb[lc], gb_c = gb[lr], ga_d = ga[lc];
    const double dl = lr == lc ? 1.0 : 0.0;
    const double bmb = 2.0 / 3.0 * ga_c * gb_d * (dl - 1.0 / 3.0) + (dl * (gab - 2.0 * ga_c * gb_c) + ga_d * gb_c) / 3.0;
    const double kcd = w * bmb + wnn * Gq[3 * lac + lr] * Gq[3 * b + lc];
    // kr[c][c'] = sum_d Ke[c][d] Rj[c'][d] for c = lr, c' = lc
    const double k0 = __shfl(kcd, src_row, 64), k1 = __shfl(kcd, src_row + 1, 64), k2 = __shfl(kcd, src_row + 2, 64);
    const double kr = k0 * rl0 + k1 * rl1 + k2 * rl2;
    // blk[r][c'] = sum_c Ri[r][c] kr[c][c'] for r = lr, c' = lc
    const double t0 = __shfl(kr, src_col, 64), t1 = __shfl(kr, src_col + 3, 64), t2 = __shfl(kr, src_col + 6, 64);
    if (pos >= 0) {
      const double* Ri = axn + 9 * pos;
      double sum = 0.0;
      sum += Ri[3 * lr] * t0;
      sum += Ri[3 * lr + 1] * t1;
      sum += Ri[3 * lr + 2] * t2;
      acc[lc * d3 + 3 * pos + lr] += sum;
    }
    wave_sync_lds();
  }
  // friction on the free axes of contact nodes (:1175-1214)
  const int nf = nfix[j];
  const int ct = contact[j];
  double fk[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  bool have_fk = false;
  if (P.friction > 0 && ct != 0 && nf < 3) {
    double c = P.friction * kfric[j] * afric[j] * (2 / M_PI);
    if (c > 0) {
      // die velocity in local axes: top die moves with (0, 0, -ram)
      double vd[3];
      for (int qq = 0; qq < 3; qq++) vd[qq] = ct == 2 ? -P.ram * Rj[3 * qq + 2] : 0.0;
      // loops over all three axes with the fixed ones skipped: every array index is a compile-time constant, so the
      // small arrays stay in registers
      double s[3] = {0, 0, 0};
      double s2 = 0;
#pragma unroll
      for (int qq = 0; qq < 3; qq++)
        if (qq >= nf) { s[qq] = v[3 * j + qq] - vd[qq]; s2 += s[qq] * s[qq]; }
      double sl = sqrt(s2), z = sl / P.u0, at = atan(z);
      en += c * (sl * at - 0.5 * P.u0 * log(1 + z * z));
      pw += c * at * sl;
      double gs = sl > 1e-12 * P.u0 ? at / sl : 1 / P.u0;
#pragma unroll
      for (int qq = 0; qq < 3; qq++)
        if (qq >= nf) rj[qq] += c * gs * s[qq];
      if (matrix) {
        double tang = c / (P.u0 * (1 + z * z));
#pragma unroll
        for (int qq = 0; qq < 3; qq++)
#pragma unroll
          for (int p = 0; p < 3; p++) {
            if (qq < nf || p < nf) continue;
            double kqp;
            if (newton && sl > 1e-12 * P.u0) kqp = c * gs * ((qq == p ? 1 : 0) - s[qq] * s[p] / s2) + tang * s[qq] * s[p] / s2;
            else kqp = qq == p ? c * gs : 0;
            fk[qq][p] = kqp;
          }
        have_fk = true;
      }
    }
  }
  if (lane == 0) { n_energy[j] = en; n_power[j] = pw; }
  if (mode == 2) return;
  if (matrix) {
    const int pj = bsearch_lds(nbr, deg, j);
    if (have_fk && lane < 9 && pj >= 0) {
      const int qq = lane / 3, p = lane % 3;
      double f = 0.0;
#pragma unroll
      for (int a2 = 0; a2 < 3; a2++)
#pragma unroll
        for (int b2 = 0; b2 < 3; b2++)
          if (lane == 3 * a2 + b2) f = fk[a2][b2];
      acc[p * d3 + 3 * pj + qq] += f;
    }
    wave_sync_lds();
    double dmax = 0.0;
    if (mode == 1) {
      // identity rows/columns for fixed DOFs, largest free diagonal. The fixed-DOF counts of the rows a lane holds
      // (entries lane + 64 k of every column) are gathered at once; nodes with more column entries than the batch
      // covers take the plain loop
      const bool batched = d3 <= 64 * NFIX_BATCH;
      int nfix_l[NFIX_BATCH];
#pragma unroll
      for (int k = 0; k < NFIX_BATCH; k++) {
        const int q = lane + 64 * k;
        nfix_l[k] = (batched && q < d3) ? nfix[nbr[q / 3]] : 0;
      }
      auto fix_entry = [&](int c, int q, int nfi) {
        const int pos = q / 3, rr = q % 3;
        const int i = nbr[pos];
        const bool diag = (i == j) && (rr == c);
        double vv = acc[c * d3 + q];
        if (c < nf || rr < nfi) vv = diag ? 1.0 : 0.0;
        else if (diag) dmax = fmax(dmax, vv);
        acc[c * d3 + q] = vv;
      };
      for (int c = 0; c < 3; c++) {
        if (batched) {
#pragma unroll
          for (int k = 0; k < NFIX_BATCH; k++)
            if (lane + 64 * k < d3) fix_entry(c, lane + 64 * k, nfix_l[k]);
        } else {
          for (int q = lane; q < d3; q += 64) fix_entry(c, q, nfix[nbr[q / 3]]);
        }
      }
      for (int off = 32; off > 0; off >>= 1) dmax = fmax(dmax, __shfl_down(dmax, off, 64));
      if (lane == 0) atomicMax(maxdiag, (unsigned long long)__double_as_longlong(fmax(dmax, 0.0)));
    }
    for (int c = 0; c < 3; c++) {
      double* dst = val + colPtr[3 * j + c];
      for (int q = lane; q < d3; q += 64) dst[q] = acc[c * d3 + q];
    }
  }
  if (lane < 3) {
    double rv = lane == 0 ? rj[0] : (lane == 1 ? rj[1] : rj[2]);
    if (mode == 1) rv = lane < nf ? -0.0 : -rv;
    r[3 * j + lane] = rv;
  }
}

// Residual-only assembly (line-search trials, final evaluation, start judging): one wave per node, the K_vol row dot
// with one index load and one gather of v for the node's three columns, the elements around the node one per lane
// (all their loads in flight) summed over the wave in a fixed order, friction, then energy, power and r as k_node
// writes them for the same mode (0 raw residual, 1 BC + negated residual, 2 energy and power only).
__global__ __launch_bounds__(256) void k_node_res(int nn, ifea_step_params P, int mode, const int* __restrict__ tets,
                                                  const int* __restrict__ nts, const int* __restrict__ ntl,
                                                  const int* __restrict__ colPtr, const int* __restrict__ rowIdx,
                                                  const double* __restrict__ ax, const int* __restrict__ nfix,
                                                  const int* __restrict__ contact, const double* __restrict__ afric,
                                                  const double* __restrict__ kfric, const double* __restrict__ kvol,
                                                  const double* __restrict__ v, const double* __restrict__ wm,
                                                  const double* __restrict__ G, double* __restrict__ r,
                                                  double* __restrict__ n_energy, double* __restrict__ n_power) {
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int j = blockIdx.x * NODE_WAVES + wv;
  if (j >= nn) return;
  const int p0 = colPtr[3 * j], p1 = colPtr[3 * j + 1], p2 = colPtr[3 * j + 2];
  const int d3 = p1 - p0;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int q = lane; q < d3; q += 64) {
    const double vv = v[rowIdx[p0 + q]];
    s0 += kvol[p0 + q] * vv;
    s1 += kvol[p1 + q] * vv;
    s2 += kvol[p2 + q] * vv;
  }
  double rj[3] = {wsum(s0), wsum(s1), wsum(s2)};
  double en = 0.5 * v[3 * j] * rj[0] + 0.5 * v[3 * j + 1] * rj[1] + 0.5 * v[3 * j + 2] * rj[2];
  double pw = 2.0 * en;
  const double* Rj = ax + 9 * (long long)j;
  double c3[3] = {0.0, 0.0, 0.0};
  for (int q = nts[j] + lane; q < nts[j + 1]; q += 64) {
    const int e = ntl[q];
    const int* t = tets + 4 * e;
    const int b = t[0] == j ? 0 : (t[1] == j ? 1 : (t[2] == j ? 2 : 3));
    const double w = wm[e];
    const double* Gb = G + 12 * (long long)e + 3 * b;
    const double reb0 = w * Gb[0], reb1 = w * Gb[1], reb2 = w * Gb[2];
#pragma unroll
    for (int c = 0; c < 3; c++) c3[c] += reb0 * Rj[3 * c] + reb1 * Rj[3 * c + 1] + reb2 * Rj[3 * c + 2];
  }
#pragma unroll
  for (int c = 0; c < 3; c++) rj[c] += wsum(c3[c]);
  const int nf = nfix[j];
  node_friction(P, j, nf, contact[j], Rj, v, afric, kfric, en, pw, rj);
  if (lane == 0) { n_energy[j] = en; n_power[j] = pw; }
  if (mode == 2) return;
  if (lane < 3) {
    double rv = lane == 0 ? rj[0] : (lane == 1 ? rj[1] : rj[2]);
    if (mode == 1) rv = lane < nf ? -0.0 : -rv;
    r[3 * j + lane] = rv;
  }
}

__global__ void k_reg(int nn, const int* __restrict__ nfix, const int* __restrict__ diagpos,
                      const unsigned long long* __restrict__ maxdiag, double* __restrict__ val) {
  int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= 3 * nn) return;
  if (d % 3 < nfix[d / 3]) return;
  double reg = 1e-10 * __longlong_as_double((long long)*maxdiag);
  val[diagpos[d]] += reg;
}

__global__ void k_trial(long long n, const double* __restrict__ v, double beta, const double* __restrict__ dv,
                        double* __restrict__ vt) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) vt[i] = v[i] + beta * dv[i];
}

// FinalEvaluation (:1402-1451) node part
__global__ void k_final_nodes(int nn, const double* __restrict__ ax, const double* __restrict__ v,
                              const double* __restrict__ r, const int* __restrict__ contact,
                              const int* __restrict__ die_axis, double* __restrict__ vg, double* __restrict__ rz,
                              double* __restrict__ comp, unsigned long long* __restrict__ maxtb) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  const double* R = ax + 9 * (long long)n;
  double l0 = v[3 * n], l1 = v[3 * n + 1], l2 = v[3 * n + 2];
  for (int c = 0; c < 3; c++) vg[3 * n + c] = l0 * R[c] + l1 * R[3 + c] + l2 * R[6 + c];
  double rzv = 0.0, cp = 0.0;
  int ct = contact[n];
  if (ct != 0) {
    double q0 = r[3 * n], q1 = r[3 * n + 1], q2 = r[3 * n + 2];
    double gz = q0 * R[2] + q1 * R[5] + q2 * R[8];
    if (ct == 2) rzv = -gz;
    int da = die_axis[n];
    if (da >= 0) {
      cp = r[3 * n + da];
      atomicMax(maxtb + (ct == 2 ? 0 : 1), (unsigned long long)__double_as_longlong(fmax(cp, 0.0)));
    }
  }
  rz[n] = rzv;
  comp[n] = cp;
}

// TetMesh.Quality (TetMesh.cs:107-112) of every tet at the node positions x: 6 sqrt(2) V / L_rms^3 with L_rms^2 the
// mean squared edge length (1 for a regular tet, <= 0 inverted, -1 for a degenerate one). The smallest index of a tet
// at or below qmin goes to *first by an integer atomic minimum (the same index whatever the order of the lanes).
__global__ void k_quality(int ne, const int* __restrict__ tets, const double* __restrict__ x, double qmin,
                          double* __restrict__ q, int* __restrict__ first) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  double p[4][3];
#pragma unroll
  for (int a = 0; a < 4; a++) {
    const int n = tets[4 * e + a];
#pragma unroll
    for (int c = 0; c < 3; c++) p[a][c] = x[3 * n + c];
  }
  auto d2 = [&](int a, int b) {
    const double dx = p[b][0] - p[a][0], dy = p[b][1] - p[a][1], dz = p[b][2] - p[a][2];
    return dx * dx + dy * dy + dz * dz;
  };
  const double l2 = (d2(0, 1) + d2(0, 2) + d2(0, 3) + d2(1, 2) + d2(1, 3) + d2(2, 3)) / 6.0;
  double u[3], v[3], w[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    u[c] = p[1][c] - p[0][c];
    v[c] = p[2][c] - p[0][c];
    w[c] = p[3][c] - p[0][c];
  }
  const double vol = (u[0] * (v[1] * w[2] - v[2] * w[1]) + u[1] * (v[2] * w[0] - v[0] * w[2]) +
                      u[2] * (v[0] * w[1] - v[1] * w[0])) / 6.0;
  const double qe = l2 > 0.0 ? 6.0 * sqrt(2.0) * vol / (l2 * sqrt(l2)) : -1.0;
  if (q) q[e] = qe;
  if (qe <= qmin) atomicMin(first, e);
}

__global__ void k_advance(int ne, double dt, const double* __restrict__ rate, double* __restrict__ eps) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < ne) eps[e] += rate[e] * dt;
}

// FinalEvaluation's element mean stress (:1440-1450): the nodal (ANP) volumetric rate kPen g_n . vg from the global
// velocities of the converged field, then the mean of an element's four nodes
__global__ void k_nodal_pressure(int nn, double kpen, const int* __restrict__ ps, const int* __restrict__ pn,
                                 const double* __restrict__ gcoef, const double* __restrict__ vg,
                                 double* __restrict__ nodal) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nn) return;
  double ev = 0.0;
  for (int q = ps[n]; q < ps[n + 1]; q++) {
    const int m = pn[q];
    const double* c = gcoef + 3 * (long long)q;
    ev += c[0] * vg[3 * m] + c[1] * vg[3 * m + 1] + c[2] * vg[3 * m + 2];
  }
  nodal[n] = kpen * ev;
}

__global__ void k_el_mean(int ne, const int* __restrict__ tets, const double* __restrict__ nodal,
                          double* __restrict__ mean) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int* t = tets + 4 * (long long)e;
  mean[e] = 0.25 * (nodal[t[0]] + nodal[t[1]] + nodal[t[2]] + nodal[t[3]]);
}

// Largest eigenvalue of a symmetric 3x3 tensor by the trigonometric solution of its characteristic cubic
// (Principal.Max3, Solver/Principal.cs:7-19)
__device__ double principal_max3(double xx, double yy, double zz, double xy, double yz, double zx) {
  const double p1 = xy * xy + yz * yz + zx * zx;
  const double q = (xx + yy + zz) / 3;
  if (p1 < 1e-30 * (xx * xx + yy * yy + zz * zz + 1e-300)) return fmax(xx, fmax(yy, zz));
  const double p2 = (xx - q) * (xx - q) + (yy - q) * (yy - q) + (zz - q) * (zz - q) + 2 * p1;
  const double p = sqrt(p2 / 6);
  const double bxx = (xx - q) / p, byy = (yy - q) / p, bzz = (zz - q) / p, bxy = xy / p, byz = yz / p, bzx = zx / p;
  const double det = bxx * (byy * bzz - byz * byz) - bxy * (bxy * bzz - byz * bzx) + bzx * (bxy * byz - byy * bzx);
  const double r = fmin(1.0, fmax(-1.0, det / 2));
  return q + 2 * p * cos(acos(r) / 3);
}

// AccumulateDamage (:1844-1855): normalised Cockcroft-Latham D += max(sigma_1, 0) / sigma_bar * rate * dt with
// sigma = s + sigma_m I from the converged deviatoric stress and the element mean stress; tension only
__global__ void k_damage(int ne, double dt, double rate_cut, const double* __restrict__ sig,
                         const double* __restrict__ rate, const double* __restrict__ dev, const double* __restrict__ mean,
                         double* __restrict__ dmg) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const double s = sig[e], re = rate[e];
  if (!(s > 0) || re <= rate_cut) return;
  const double* d = dev + 6 * (long long)e;
  const double m = mean[e];
  const double s1 = principal_max3(d[0] + m, d[1] + m, d[2] + m, d[3], d[4], d[5]);
  if (s1 > 0) dmg[e] += s1 / s * re * dt;
}

// ------------------------------------------------------------------------------- host
StepDevice::StepDevice(int nn, int ne, const int* tets, const int* colPtr, const int* rowIdx, int nf, const int* faces,
                       const int* face_owner)
    : nn_(nn), ne_(ne), ndof_(3 * nn) {
  if (nn <= 0 || ne <= 0 || !tets || !colPtr || !rowIdx) throw Error(-1, "empty mesh");
  IFEA_HIP(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
  vec_.reset(new Vec(s_));
  spd_.reset(new SpdDevice(s_));
  nnz_ = colPtr[ndof_];
  h_colPtr_.assign(colPtr, colPtr + ndof_ + 1);
  h_rowIdx_.assign(rowIdx, rowIdx + nnz_);
  // node -> tets (ascending), TetMesh.cs:18-29
  std::vector<int> nts(nn + 1, 0), ntl(4 * (size_t)ne);
  for (size_t k = 0; k < 4 * (size_t)ne; k++) {
    if (tets[k] < 0 || tets[k] >= nn) throw Error(-1, "tet node index out of range");
    nts[tets[k] + 1]++;
  }
  for (int n = 0; n < nn; n++) nts[n + 1] += nts[n];
  {
    std::vector<int> fill(nts.begin(), nts.end() - 1);
    for (int e = 0; e < ne; e++) for (int a = 0; a < 4; a++) ntl[fill[tets[4 * e + a]]++] = e;
  }
  // neighbour lists from the pattern: column 3j lists rows 3i, 3i+1, 3i+2 for i in nbr[j]
  std::vector<int> nbs(nn + 1, 0), nbv;
  std::vector<int> diagpos(ndof_);
  for (int j = 0; j < nn; j++) {
    int p0 = colPtr[3 * j], p1 = colPtr[3 * j + 1];
    if ((p1 - p0) % 3) throw Error(-1, "pattern is not made of 3x3 node blocks");
    int deg = (p1 - p0) / 3;
    if (deg > MAX_DEG) throw Error(-1, "node with more than 192 pattern neighbours");
    max_deg_ = std::max(max_deg_, deg);
    int pj = -1;
    for (int q = 0; q < deg; q++) {
      int i = rowIdx[p0 + 3 * q] / 3;
      if (i == j) pj = q;
      nbv.push_back(i);
    }
    if (pj < 0) throw Error(-1, "pattern misses a diagonal block");
    for (int c = 0; c < 3; c++) {
      if (colPtr[3 * j + c + 1] - colPtr[3 * j + c] != 3 * deg) throw Error(-1, "pattern is not made of 3x3 node blocks");
      diagpos[3 * j + c] = colPtr[3 * j + c] + 3 * pj + c;
    }
    nbs[j + 1] = (int)nbv.size();
  }
  // every pair of nodes of an element must be coupled in the pattern (the assembly gathers rely on it)
  for (int e = 0; e < ne; e++)
    for (int a = 0; a < 4; a++)
      for (int b = 0; b < 4; b++) {
        int i = tets[4 * e + a], j = tets[4 * e + b];
        if (!std::binary_search(nbv.begin() + nbs[j], nbv.begin() + nbs[j + 1], i))
          throw Error(-1, "pattern does not couple the nodes of element " + std::to_string(e));
      }
  // one-ring patches (sorted)
  std::vector<int> ps(nn + 1, 0), pnv;
  std::vector<int> tmp;
  for (int n = 0; n < nn; n++) {
    tmp.clear();
    for (int q = nts[n]; q < nts[n + 1]; q++) for (int a = 0; a < 4; a++) tmp.push_back(tets[4 * ntl[q] + a]);
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    if ((int)tmp.size() > MAX_PATCH) throw Error(-1, "node patch larger than 64 nodes");
    max_patch_ = std::max(max_patch_, (int)tmp.size());
    pnv.insert(pnv.end(), tmp.begin(), tmp.end());
    ps[n + 1] = (int)pnv.size();
  }
  // faces of each node in face order
  std::vector<int> nfs(nn + 1, 0), nfl;
  for (int f = 0; f < nf; f++) for (int a = 0; a < 3; a++) nfs[faces[3 * f + a] + 1]++;
  for (int n = 0; n < nn; n++) nfs[n + 1] += nfs[n];
  nfl.resize(nfs[nn]);
  {
    std::vector<int> fill(nfs.begin(), nfs.end() - 1);
    for (int f = 0; f < nf; f++) for (int a = 0; a < 3; a++) nfl[fill[faces[3 * f + a]]++] = f;
  }
  tets_.upload(tets, 4 * (size_t)ne, s_);
  for (int n = 0; n < nn; n++) max_elem_ = std::max(max_elem_, nts[n + 1] - nts[n]);
  {
    // the tangent assembly keeps a node's columns, its neighbours' frames and its elements in LDS
    const size_t lds = node_tan_lds_bytes(max_deg_, max_elem_);
    if (lds > 160 * 1024) throw Error(-1, "a node has too many neighbours and elements for the assembly kernel");
    if (lds > 64 * 1024)
      IFEA_HIP(hipFuncSetAttribute((const void*)k_node, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  }
  nt_start_.upload(nts, s_);
  nt_list_.upload(ntl, s_);
  colPtr_.upload(colPtr, ndof_ + 1, s_);
  rowIdx_.upload(rowIdx, nnz_, s_);
  nbr_start_.upload(nbs, s_);
  nbr_.upload(nbv, s_);
  p_start_.upload(ps, s_);
  p_node_.upload(pnv, s_);
  diagpos_.upload(diagpos, s_);
  nf_start_.upload(nfs, s_);
  if (nfl.empty()) nf_list_.alloc(1); else nf_list_.upload(nfl, s_);
  if (nf) { faces_.upload(faces, 3 * (size_t)nf, s_); face_owner_.upload(face_owner, nf, s_); }
  else { faces_.alloc(3); face_owner_.alloc(1); }
  x_.alloc(ndof_); ax_.alloc(9 * (size_t)nn); fixed_val_.alloc(ndof_);
  grad_.alloc(12 * (size_t)ne); vol_.alloc(ne); vnode_.alloc(nn); gcoef_.alloc(3 * pnv.size());
  kvol_.alloc(nnz_); afric_.alloc(nn); kfric_.alloc(nn);
  nfix_.alloc(nn); contact_.alloc(nn); bad_.alloc(1);
  eps_.alloc(ne); el_sig_.alloc(ne); el_rate_.alloc(ne); el_dev_.alloc(6 * (size_t)ne); el_mean_.alloc(ne);
  dmg_.alloc(ne); nodal_.alloc(nn);
  IFEA_HIP(hipMemsetAsync(eps_.p, 0, ne * sizeof(double), s_));
  IFEA_HIP(hipMemsetAsync(el_dev_.p, 0, 6 * (size_t)ne * sizeof(double), s_));
  IFEA_HIP(hipMemsetAsync(el_mean_.p, 0, ne * sizeof(double), s_));
  IFEA_HIP(hipMemsetAsync(dmg_.p, 0, ne * sizeof(double), s_));
  IFEA_HIP(hipMemsetAsync(el_sig_.p, 0, ne * sizeof(double), s_));
  IFEA_HIP(hipMemsetAsync(el_rate_.p, 0, ne * sizeof(double), s_));
  v_.alloc(ndof_); r_.alloc(ndof_); dv_.alloc(ndof_); vtrial_.alloc(ndof_); val_.alloc(nnz_);
  e_wm_.alloc(ne); e_wn_.alloc(ne); e_g_.alloc(12 * (size_t)ne); e_energy_.alloc(ne); e_power_.alloc(ne);
  n_energy_.alloc(nn); n_power_.alloc(nn);
  pr_.alloc(ndof_); pz_.alloc(ndof_); pp_.alloc(ndof_); pq_.alloc(ndof_); tmp_.alloc(3 * (size_t)nn);
  maxdiag_.alloc(4);
  IFEA_HIP(hipMemsetAsync(v_.p, 0, ndof_ * sizeof(double), s_));
  IFEA_HIP(hipStreamSynchronize(s_));
}

StepDevice::~StepDevice() {
  spd_.reset();
  vec_.reset();
  if (s_) (void)hipStreamDestroy(s_);
}

void StepDevice::set_option(int key, double value) {
  if (key == IFEA_STEP_OPT_FIRST_SOLVE) { opt_first_solve_ = value != 0 ? 1 : 0; return; }
  throw Error(-1, "unknown step option key " + std::to_string(key));
}

void StepDevice::set_perm(const int* perm) {
  if (perm) user_perm_.assign(perm, perm + ndof_);
  else user_perm_.clear();
  need_analyze_ = true;
  have_factor_ = false;
}

void StepDevice::prepare(const double* x, const double* ax, const int* nfix, const double* fixed_val,
                         const int* contact, const double* vguess) {
  if (!have_params_) throw Error(-5, "parameters must be set before PrepareStep");
  ax_.upload(ax, 9 * (size_t)nn_, s_);
  nfix_.upload(nfix, nn_, s_);
  fixed_val_.upload(fixed_val, ndof_, s_);
  contact_.upload(contact, nn_, s_);
  device_frames_ = false;
  prepare_rest(x, vguess);
}

// PrepareStep after the frames: element geometry, ANP, friction, warm start
void StepDevice::prepare_rest(const double* x, const double* vguess) {
  if (!have_params_) throw Error(-5, "parameters must be set before PrepareStep");
  x_.upload(x, ndof_, s_);
  if (need_analyze_) {
    h_coords_.resize(3 * (size_t)ndof_);
    for (int n = 0; n < nn_; n++)
      for (int c = 0; c < 3; c++)
        for (int a = 0; a < 3; a++) h_coords_[3 * (3 * (size_t)n + c) + a] = x[3 * n + a];
  }
  IFEA_HIP(hipMemcpyAsync(tmp_.p, vguess, ndof_ * sizeof(double), hipMemcpyHostToDevice, s_));
  int big = 0x7fffffff;
  IFEA_HIP(hipMemcpyAsync(bad_.p, &big, sizeof(int), hipMemcpyHostToDevice, s_));
  hipLaunchKernelGGL(k_geom, dim3(grid_for(ne_, 256)), dim3(256), 0, s_, ne_, tets_.p, x_.p, grad_.p, vol_.p, bad_.p);
  int hb = 0;
  IFEA_HIP(hipMemcpyAsync(&hb, bad_.p, sizeof(int), hipMemcpyDeviceToHost, s_));
  IFEA_HIP(hipStreamSynchronize(s_));
  if (hb != 0x7fffffff) throw Error(-1, "Tetrahedron " + std::to_string(hb) + " is inverted");
  hipLaunchKernelGGL(k_vnode, dim3(grid_for(nn_, 256)), dim3(256), 0, s_, nn_, nt_start_.p, nt_list_.p, vol_.p, vnode_.p);
  hipLaunchKernelGGL(k_gcoef, dim3(grid_for(nn_, 128)), dim3(128), 0, s_, nn_, tets_.p, nt_start_.p, nt_list_.p,
                     p_start_.p, p_node_.p, vol_.p, vnode_.p, grad_.p, gcoef_.p);
  hipLaunchKernelGGL(k_kvol, dim3(grid_for(nn_, NODE_WAVES)), dim3(256), node_lds_bytes(max_deg_), s_, max_deg_, nn_,
                     P_.kpen, colPtr_.p, nbr_start_.p, nbr_.p, p_start_.p, p_node_.p, ax_.p, vnode_.p, gcoef_.p,
                     kvol_.p);
  hipLaunchKernelGGL(k_fric, dim3(grid_for(nn_, 256)), dim3(256), 0, s_, nn_, nf_start_.p, nf_list_.p, faces_.p,
                     face_owner_.p, x_.p, el_sig_.p, afric_.p, kfric_.p);
  hipLaunchKernelGGL(k_vguess, dim3(grid_for(nn_, 256)), dim3(256), 0, s_, nn_, ax_.p, nfix_.p, fixed_val_.p, tmp_.p,
                     v_.p);
  IFEA_HIP(hipGetLastError());
  prepared_ = true;
}


void StepDevice::set_geometry(const unsigned char* on_x0, const unsigned char* on_y0, int bottom_kind,
                              const double* bottom_p, int top_kind, const double* top_p) {
  if (!on_x0 || !on_y0 || !bottom_p || !top_p) throw Error(-1, "invalid geometry arguments");
  if (bottom_kind < 0 || bottom_kind > 2 || top_kind < 0 || top_kind > 2) throw Error(-1, "unknown die kind");
  on_x0_.upload(on_x0, nn_, s_);
  on_y0_.upload(on_y0, nn_, s_);
  bottom_.kind = bottom_kind;
  bottom_.side = 0;
  top_.kind = top_kind;
  top_.side = 1;
  for (int k = 0; k < 8; k++) { bottom_.p[k] = bottom_p[k]; top_.p[k] = top_p[k]; }
  die_axis_.alloc(nn_);
  IFEA_HIP(hipStreamSynchronize(s_));
  have_geometry_ = true;
}

void StepDevice::prepare_device(const double* x, const int* contact, const double* vguess) {
  if (!have_geometry_) throw Error(-5, "the symmetry flags and dies must be set before PrepareStep on the device");
  contact_.upload(contact, nn_, s_);
  xq_.upload(x, ndof_, s_);
  hipLaunchKernelGGL(k_frames, dim3(grid_for(nn_, 256)), dim3(256), 0, s_, nn_, bottom_, top_, P_.ram, xq_.p,
                     on_x0_.p, on_y0_.p, contact_.p, ax_.p, nfix_.p, die_axis_.p, fixed_val_.p);
  IFEA_HIP(hipGetLastError());
  device_frames_ = true;
  prepare_rest(x, vguess);
}

void StepDevice::load_guess(const double* vguess) {
  if (!prepared_) throw Error(-5, "PrepareStep must be called first");
  IFEA_HIP(hipMemcpyAsync(tmp_.p, vguess, ndof_ * sizeof(double), hipMemcpyHostToDevice, s_));
  hipLaunchKernelGGL(k_vguess, dim3(grid_for(nn_, 256)), dim3(256), 0, s_, nn_, ax_.p, nfix_.p, fixed_val_.p, tmp_.p,
                     v_.p);
  IFEA_HIP(hipGetLastError());
  IFEA_HIP(hipStreamSynchronize(s_));
}

void StepDevice::get_frames(double* ax, int* nfix, int* die_axis, double* fixed_val) {
  if (!prepared_) throw Error(-5, "PrepareStep must be called first");
  if (ax) IFEA_HIP(hipMemcpyAsync(ax, ax_.p, 9 * (size_t)nn_ * sizeof(double), hipMemcpyDeviceToHost, s_));
  if (nfix) IFEA_HIP(hipMemcpyAsync(nfix, nfix_.p, nn_ * sizeof(int), hipMemcpyDeviceToHost, s_));
  if (die_axis) {
    if (!device_frames_) throw Error(-5, "the die axes are the caller's when the frames come from the host");
    IFEA_HIP(hipMemcpyAsync(die_axis, die_axis_.p, nn_ * sizeof(int), hipMemcpyDeviceToHost, s_));
  }
  if (fixed_val) IFEA_HIP(hipMemcpyAsync(fixed_val, fixed_val_.p, ndof_ * sizeof(double), hipMemcpyDeviceToHost, s_));
  IFEA_HIP(hipStreamSynchronize(s_));
}

void StepDevice::set_element_state(const double* eps, const double* sigma) {
  if (eps) IFEA_HIP(hipMemcpyAsync(eps_.p, eps, ne_ * sizeof(double), hipMemcpyHostToDevice, s_));
  if (sigma) IFEA_HIP(hipMemcpyAsync(el_sig_.p, sigma, ne_ * sizeof(double), hipMemcpyHostToDevice, s_));
  IFEA_HIP(hipStreamSynchronize(s_));
}

double StepDevice::assemble(bool matrix, bool newton, const double* d_v, int mode, double* power) {
  hipLaunchKernelGGL(k_elem, dim3(grid_for(ne_, 256)), dim3(256), 0, s_, ne_, P_, newton ? 1 : 0, tets_.p, ax_.p, d_v,
                     grad_.p, vol_.p, eps_.p, e_wm_.p, e_wn_.p, e_g_.p, e_energy_.p, e_power_.p, el_sig_.p,
                     el_rate_.p, el_dev_.p);
  if (mode == 1) IFEA_HIP(hipMemsetAsync(maxdiag_.p, 0, sizeof(unsigned long long), s_));
  if (matrix)
    hipLaunchKernelGGL(k_node, dim3(grid_for(nn_, TAN_WAVES)), dim3(64 * TAN_WAVES),
                       node_tan_lds_bytes(max_deg_, max_elem_), s_, max_deg_, max_elem_, nn_, P_, 1, newton ? 1 : 0,
                       mode, tets_.p, nt_start_.p, nt_list_.p, colPtr_.p, rowIdx_.p, nbr_start_.p, nbr_.p, ax_.p,
                       nfix_.p, contact_.p, afric_.p, kfric_.p, kvol_.p, d_v, grad_.p, e_wm_.p, e_wn_.p, e_g_.p, r_.p,
                       val_.p, n_energy_.p, n_power_.p, maxdiag_.p);
  else
    hipLaunchKernelGGL(k_node_res, dim3(grid_for(nn_, NODE_WAVES)), dim3(256), 0, s_, nn_, P_, mode, tets_.p,
                       nt_start_.p, nt_list_.p, colPtr_.p, rowIdx_.p, ax_.p, nfix_.p, contact_.p, afric_.p, kfric_.p,
                       kvol_.p, d_v, e_wm_.p, e_g_.p, r_.p, n_energy_.p, n_power_.p);
  if (matrix && mode == 1)
    hipLaunchKernelGGL(k_reg, dim3(grid_for(ndof_, 256)), dim3(256), 0, s_, nn_, nfix_.p, diagpos_.p, maxdiag_.p,
                       val_.p);
  vec_->dot(n_energy_.p, nullptr, nn_, S_E_NODE);
  vec_->dot(e_energy_.p, nullptr, ne_, S_E_ELEM);
  vec_->dot(n_power_.p, nullptr, nn_, S_P_NODE);
  vec_->dot(e_power_.p, nullptr, ne_, S_P_ELEM);
  double sc[4];
  vec_->read(S_E_NODE, 4, sc);
  if (power) *power = sc[2] + sc[3];
  return sc[0] + sc[1];
}

void StepDevice::assemble_host(int matrix, int newton, const double* v, double* energy, double* power, double* r,
                               double* val) {
  if (!prepared_) throw Error(-5, "PrepareStep must be called first");
  if (v) IFEA_HIP(hipMemcpyAsync(v_.p, v, ndof_ * sizeof(double), hipMemcpyHostToDevice, s_));
  double pw = 0;
  double e = assemble(matrix != 0, newton != 0, v_.p, 0, &pw);
  if (energy) *energy = e;
  if (power) *power = pw;
  if (r) IFEA_HIP(hipMemcpyAsync(r, r_.p, ndof_ * sizeof(double), hipMemcpyDeviceToHost, s_));
  if (val && matrix) IFEA_HIP(hipMemcpyAsync(val, val_.p, nnz_ * sizeof(double), hipMemcpyDeviceToHost, s_));
  IFEA_HIP(hipStreamSynchronize(s_));
}

int StepDevice::refactor(int precision, ifea_step_stats* st) {
  if (need_analyze_) {
    SymbolicOptions opt;
    if (!h_coords_.empty()) opt.coords = h_coords_.data();
    spd_->analyze(ndof_, h_colPtr_.data(), h_rowIdx_.data(), user_perm_.empty() ? nullptr : user_perm_.data(), opt);
    need_analyze_ = false;
  }
  // a factorization that throws (out of memory, a device error) must not leave PCG with a half-written factor
  have_factor_ = false;
  int rc = spd_->factorize_device(val_.p, precision);
  if (st) {
    st->factorizations++;
    st->ms_factor += spd_->last_factor_ms();
  }
  return rc;
}

void StepDevice::lin_solve(const double* b, double* x, ifea_step_stats* st) {
  spd_->solve_device(b, x);
  if (st) st->solves++;
}

// Pcg (:1051-1087): preconditioner = the cached factor; returns iterations or -1.
int StepDevice::pcg(const double* b, double* x, double tol, int maxIt, ifea_step_stats* st) {
  if (st) st->pcg_calls++;
  vec_->dot(b, b, ndof_, S_DOT0);
  double bn2;
  vec_->read(S_DOT0, 1, &bn2);
  double bnorm = std::sqrt(bn2);
  if (bnorm == 0) {
    IFEA_HIP(hipMemsetAsync(x, 0, ndof_ * sizeof(double), s_));
    return 0;
  }
  // The reference tests the residual at the top of each iteration, after the preconditioner has been applied to it;
  // here the test comes right after the residual update, so the preconditioner sweep of the residual that meets the
  // tolerance (or ends the budget) is not made. Iterates, tests and return values are those of the reference.
  auto residual_test = [&]() {
    vec_->dot(pr_.p, pr_.p, ndof_, S_DOT1);
    double rn2;
    vec_->read(S_DOT1, 1, &rn2);
    if (!std::isfinite(rn2)) return -2;
    return std::sqrt(rn2) <= tol * bnorm ? 1 : 0;
  };
  lin_solve(b, x, st);
  spmv_sym(s_, ndof_, colPtr_.p, rowIdx_.p, val_.p, x, pq_.p);
  IFEA_HIP(hipMemcpyAsync(pr_.p, b, ndof_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
  axpby(s_, ndof_, -1.0, pq_.p, 1.0, pr_.p);
  int t = residual_test();
  if (t != 0) return t > 0 ? 0 : -2;
  if (maxIt <= 0) return -1;
  lin_solve(pr_.p, pz_.p, st);
  IFEA_HIP(hipMemcpyAsync(pp_.p, pz_.p, ndof_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
  vec_->dot(pr_.p, pz_.p, ndof_, S_RZ);
  for (int it = 0; it < maxIt; it++) {
    if (st) st->pcg_iterations++;
    spmv_sym(s_, ndof_, colPtr_.p, rowIdx_.p, val_.p, pp_.p, pq_.p);
    vec_->dot(pp_.p, pq_.p, ndof_, S_PQ);
    pcg_update_x_r(s_, ndof_, vec_->slots(), pp_.p, pq_.p, x, pr_.p);
    t = residual_test();
    if (t != 0) return t > 0 ? it + 1 : -2;
    if (it + 1 == maxIt) break;
    lin_solve(pr_.p, pz_.p, st);
    vec_->dot(pr_.p, pz_.p, ndof_, S_RZN);
    pcg_update_p(s_, ndof_, vec_->slots(), pz_.p, pp_.p);
  }
  return -1;
}

// SolveLinear (:965-1019) with the mixed-precision refinement slot
int StepDevice::solve_linear(double* b, double* x, ifea_step_stats* st) {
  const bool first = first_solve_;
  first_solve_ = false;
  // IFEA_STEP_OPT_FIRST_SOLVE: the first solve of an increment meets the previous increment's factor, which the
  // changed contact set and frames make a poor preconditioner; after it failed the PCG budget at the first solves of
  // two increments in a row, the first solve refactors directly, probing PCG again after 8 skipped increments.
  const bool skip_first = first && opt_first_solve_ && have_factor_ && first_fail_streak_ >= 2 && first_skips_ < 8;
  if (skip_first) {
    first_skips_++;
    if (st) st->first_solve_skips++;
  } else if (have_factor_ && pcg_skip_ > 0) {
    pcg_skip_--;
  } else if (have_factor_) {
    int its = pcg(b, x, P_.pcg_tol, P_.pcg_budget, st);
    if (first) {
      first_fail_streak_ = its >= 0 ? 0 : first_fail_streak_ + 1;
      first_skips_ = 0;
    }
    if (its >= 0) { pcg_failures_ = 0; return 0; }
    if (++pcg_failures_ >= 2) pcg_skip_ = std::min(64, 1 << (pcg_failures_ - 1));
  }
  // precision 1: FP32 factor; after an FP32 factor that is not positive definite (penalty scale beyond FP32), FP64
  // factors with an FP32 copy for the solves. Both are refined by PCG in FP64 against the FP64 matrix. When the
  // refinement of a fresh factor with FP32 sweeps stalls, the FP32 copy does not resolve this matrix: plain FP64
  // from then on (trying the copy again after every factorisation would cost a stalled PCG each time).
  int prec = P_.precision == 0 ? 0 : (fp32_sweeps_unstable_ ? 0 : (P_.precision == 1 && !fp32_unstable_ ? 1 : 2));
  int rc = refactor(prec, st);
  if (rc == -4 && prec == 1) {
    fp32_unstable_ = true;
    if (st) st->fp64_fallbacks++;
    prec = 2;
    rc = refactor(prec, st);
  }
  if (rc) { have_factor_ = false; return rc; }
  have_factor_ = true;
  if (prec != 0) {
    int its = pcg(b, x, P_.pcg_tol, std::max(P_.pcg_budget, 30), st);
    if (its >= 0) return 0;
    if (st) st->fp64_fallbacks++;
    if (prec == 2) {
      spd_->use_fp64_solves(true);
      fp32_sweeps_unstable_ = true;
    } else {
      rc = refactor(0, st);
      if (rc) { have_factor_ = false; return rc; }
    }
  }
  lin_solve(b, x, st);
  return 0;
}

// SolveVelocity (:875-948)
void StepDevice::solve_velocity(int warm, ifea_step_stats* st) {
  if (!prepared_) throw Error(-5, "PrepareStep must be called first");
  ifea_step_stats local{};
  ifea_step_stats* S = st ? st : &local;
  std::memset(S, 0, sizeof(*S));
  auto t0 = std::chrono::steady_clock::now();
  int halved = 0;
  double last_rel = INFINITY;
  int it;
  first_solve_ = true;
  const int nlim = P_.newton_limit > 0 ? P_.newton_limit : 20;
  for (it = 1; it <= P_.max_iterations; it++) {
    bool newton = (warm || it > 2) && it <= nlim && !(P_.adaptive_newton && halved >= 2);
    if (newton) S->newton_iterations++;
    double pw;
    double e0 = assemble(true, newton, v_.p, 1, &pw);
    S->assemblies++;
    // r now holds -r with fixed rows zeroed
    vec_->dot(r_.p, r_.p, ndof_, S_DOT2);
    int rc = solve_linear(r_.p, dv_.p, S);
    if (rc) throw Error(rc == -4 ? -4 : -6, "Sparse factorisation failed (matrix not positive definite)");
    vec_->dot(r_.p, dv_.p, ndof_, S_DOT3);
    double sl[2];
    vec_->read(S_DOT2, 2, sl);
    S->last_rnorm = std::sqrt(sl[0]);
    double slope = -sl[1];
    double beta = 1;
    for (int ls = 0; ls < 8; ls++) {
      hipLaunchKernelGGL(k_trial, dim3(grid_for(ndof_, 256)), dim3(256), 0, s_, (long long)ndof_, v_.p, beta, dv_.p,
                         vtrial_.p);
      double p1;
      double e1 = assemble(false, false, vtrial_.p, 2, &p1);
      S->assemblies++;
      if (e1 <= e0 + 1e-4 * beta * slope + 1e-11 * std::fabs(e0)) break;
      double curv = (e1 - e0 - slope * beta) / (beta * beta);
      beta = P_.quadratic_ls && curv > 0 ? std::min(0.7 * beta, std::max(0.1 * beta, -slope / (2 * curv))) : 0.5 * beta;
    }
    IFEA_HIP(hipMemcpyAsync(v_.p, vtrial_.p, ndof_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
    halved = newton && beta < 1 ? halved + 1 : 0;
    vec_->dot(dv_.p, dv_.p, ndof_, S_DOT0);
    vec_->dot(v_.p, v_.p, ndof_, S_DOT1);
    double nn2[2];
    vec_->read(S_DOT0, 2, nn2);
    double rel = beta * std::sqrt(nn2[0]) / std::max(std::sqrt(nn2[1]), 1e-30);
    last_rel = rel;
    if (std::isnan(rel)) throw Error(-6, "Velocity solution became NaN");
    if (rel < P_.tolerance) { S->converged = 1; break; }
  }
  S->iterations = std::min(it, P_.max_iterations);
  S->last_rel = last_rel;
  S->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void StepDevice::final_eval(const int* die_axis, double* power, double* reaction, double* vg, double* r,
                            unsigned char* release) {
  double pw = 0;
  assemble(false, false, v_.p, 0, &pw);
  if (power) *power = pw;
  DBuf<int> da;
  std::vector<int> zero(nn_, -1);
  da.upload(die_axis ? die_axis : zero.data(), nn_, s_);
  DBuf<double> vgd, rz, comp;
  vgd.alloc(ndof_);
  rz.alloc(nn_);
  comp.alloc(nn_);
  IFEA_HIP(hipMemsetAsync(maxdiag_.p, 0, 2 * sizeof(unsigned long long), s_));
  hipLaunchKernelGGL(k_final_nodes, dim3(grid_for(nn_, 256)), dim3(256), 0, s_, nn_, ax_.p, v_.p, r_.p, contact_.p,
                     da.p, vgd.p, rz.p, comp.p, maxdiag_.p);
  hipLaunchKernelGGL(k_nodal_pressure, dim3(grid_for(nn_, 256)), dim3(256), 0, s_, nn_, P_.kpen, p_start_.p,
                     p_node_.p, gcoef_.p, vgd.p, nodal_.p);
  hipLaunchKernelGGL(k_el_mean, dim3(grid_for(ne_, 256)), dim3(256), 0, s_, ne_, tets_.p, nodal_.p, el_mean_.p);
  IFEA_HIP(hipGetLastError());
  vec_->dot(rz.p, nullptr, nn_, S_REACTION);
  double rc;
  vec_->read(S_REACTION, 1, &rc);
  if (reaction) *reaction = rc;
  if (vg) IFEA_HIP(hipMemcpyAsync(vg, vgd.p, ndof_ * sizeof(double), hipMemcpyDeviceToHost, s_));
  if (r) IFEA_HIP(hipMemcpyAsync(r, r_.p, ndof_ * sizeof(double), hipMemcpyDeviceToHost, s_));
  if (release) {
    // tensile release (:1421-1439), small per-node pass on the host
    std::vector<double> cp(nn_);
    std::vector<int> ct(nn_);
    unsigned long long mx[2];
    IFEA_HIP(hipMemcpyAsync(cp.data(), comp.p, nn_ * sizeof(double), hipMemcpyDeviceToHost, s_));
    IFEA_HIP(hipMemcpyAsync(ct.data(), contact_.p, nn_ * sizeof(int), hipMemcpyDeviceToHost, s_));
    IFEA_HIP(hipMemcpyAsync(mx, maxdiag_.p, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s_));
    IFEA_HIP(hipStreamSynchronize(s_));
    double maxT, maxB;
    std::memcpy(&maxT, &mx[0], 8);
    std::memcpy(&maxB, &mx[1], 8);
    int keptB = 0, keptT = 0;
    for (int n = 0; n < nn_; n++) {
      release[n] = 0;
      if (ct[n] == 0 || !die_axis || die_axis[n] < 0) continue;
      double thr = 0.02 * (ct[n] == 2 ? maxT : maxB);
      if (cp[n] < -thr) release[n] = 1;
      else if (ct[n] == 2) keptT++;
      else keptB++;
    }
    if (keptB == 0 || keptT == 0)
      for (int n = 0; n < nn_; n++)
        if ((keptT == 0 && ct[n] == 2) || (keptB == 0 && ct[n] == 1)) release[n] = 0;
  }
  IFEA_HIP(hipStreamSynchronize(s_));
}

void StepDevice::advance(double dt, double* eps, double* sigma, double* rate) {
  hipLaunchKernelGGL(k_advance, dim3(grid_for(ne_, 256)), dim3(256), 0, s_, ne_, dt, el_rate_.p, eps_.p);
  hipLaunchKernelGGL(k_damage, dim3(grid_for(ne_, 256)), dim3(256), 0, s_, ne_, dt, P_.rate_cut, el_sig_.p,
                     el_rate_.p, el_dev_.p, el_mean_.p, dmg_.p);
  IFEA_HIP(hipGetLastError());
  if (eps) IFEA_HIP(hipMemcpyAsync(eps, eps_.p, ne_ * sizeof(double), hipMemcpyDeviceToHost, s_));
  if (sigma) IFEA_HIP(hipMemcpyAsync(sigma, el_sig_.p, ne_ * sizeof(double), hipMemcpyDeviceToHost, s_));
  if (rate) IFEA_HIP(hipMemcpyAsync(rate, el_rate_.p, ne_ * sizeof(double), hipMemcpyDeviceToHost, s_));
  IFEA_HIP(hipStreamSynchronize(s_));
}

void StepDevice::element_results(double* mean, double* damage, double* dev6) {
  if (mean) IFEA_HIP(hipMemcpyAsync(mean, el_mean_.p, ne_ * sizeof(double), hipMemcpyDeviceToHost, s_));
  if (damage) IFEA_HIP(hipMemcpyAsync(damage, dmg_.p, ne_ * sizeof(double), hipMemcpyDeviceToHost, s_));
  if (dev6) IFEA_HIP(hipMemcpyAsync(dev6, el_dev_.p, 6 * (size_t)ne_ * sizeof(double), hipMemcpyDeviceToHost, s_));
  IFEA_HIP(hipStreamSynchronize(s_));
}

void StepDevice::set_damage(const double* damage) {
  if (damage) IFEA_HIP(hipMemcpyAsync(dmg_.p, damage, ne_ * sizeof(double), hipMemcpyHostToDevice, s_));
  else IFEA_HIP(hipMemsetAsync(dmg_.p, 0, ne_ * sizeof(double), s_));
  IFEA_HIP(hipStreamSynchronize(s_));
}

int StepDevice::quality(const double* x, double qmin, double* q) {
  xq_.upload(x, 3 * (size_t)nn_, s_);
  bad_.alloc(1);
  const int big = 0x7fffffff;
  IFEA_HIP(hipMemcpyAsync(bad_.p, &big, sizeof(int), hipMemcpyHostToDevice, s_));
  if (q) qual_.alloc(ne_);
  hipLaunchKernelGGL(k_quality, dim3(grid_for(ne_, 256)), dim3(256), 0, s_, ne_, tets_.p, xq_.p, qmin,
                     q ? qual_.p : nullptr, bad_.p);
  IFEA_HIP(hipGetLastError());
  int first = big;
  IFEA_HIP(hipMemcpyAsync(&first, bad_.p, sizeof(int), hipMemcpyDeviceToHost, s_));
  if (q) IFEA_HIP(hipMemcpyAsync(q, qual_.p, ne_ * sizeof(double), hipMemcpyDeviceToHost, s_));
  IFEA_HIP(hipStreamSynchronize(s_));
  return first == big ? -1 : first;
}

// ------------------------------------------------------------------------------- quality check without allocation
// The collapse check of TryUpdate for callers that cannot allocate device memory: the caller's tets and positions are
// staged, the first-index counter kept and the per-tet values returned through one file-scope device array. Each GPU
// holds its own copy of the array (NOALLOC_SCRATCH_BYTES) from module load until the process ends; every call on a
// GPU shares it, so calls that overlap on one GPU race.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
constexpr int MAX_DEVICES = 64;
unsigned char* g_scratch_ptr[MAX_DEVICES] = {nullptr};

// Device address of this GPU's scratch array, looked up once per device with that device current. The array's name
// is only a host-side stand-in: host code uses this address, never the name.
unsigned char* noalloc_scratch() {
  int dev = 0;
  IFEA_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) throw Error(-6, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    IFEA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error(-6, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[dev] = (unsigned char*)p;
  }
  return g_scratch_ptr[dev];
}

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
}  // namespace

__global__ void k_set_int(int* __restrict__ p, int v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *p = v;
}

int quality_noalloc(int nn, int ne, const int* tets, const double* x, double qmin, double* q) {
  if (nn <= 0 || ne <= 0 || !tets || !x) throw Error(-1, "empty mesh");
  // layout: [0, 256) the counter; tets (4 ne ints); positions (3 nn doubles); per-tet values (ne doubles, if asked)
  const size_t o_t = 256;
  const size_t o_x = o_t + align256(4 * (size_t)ne * sizeof(int));
  const size_t o_q = o_x + align256(3 * (size_t)nn * sizeof(double));
  const size_t end = o_q + (q ? align256((size_t)ne * sizeof(double)) : 0);
  if (end > NOALLOC_SCRATCH_BYTES)
    throw Error(-1, "mesh of " + std::to_string(nn) + " nodes and " + std::to_string(ne) + " tets needs " +
                        std::to_string(end) + " bytes of scratch; the allocation-free quality check holds at most " +
                        std::to_string(NOALLOC_SCRATCH_BYTES));
  for (size_t k = 0; k < 4 * (size_t)ne; k++)
    if (tets[k] < 0 || tets[k] >= nn) throw Error(-1, "tet node index out of range");
  unsigned char* base = noalloc_scratch();
  int* first = (int*)base;
  int* d_tets = (int*)(base + o_t);
  double* d_x = (double*)(base + o_x);
  double* d_q = q ? (double*)(base + o_q) : nullptr;
  const int big = 0x7fffffff;
  hipStream_t s = 0;
  // the counter starts at "none" for every call (a fill kernel, so the sequence can also be captured)
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, s, first, big);
  IFEA_HIP(hipMemcpyAsync(d_tets, tets, 4 * (size_t)ne * sizeof(int), hipMemcpyHostToDevice, s));
  IFEA_HIP(hipMemcpyAsync(d_x, x, 3 * (size_t)nn * sizeof(double), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_quality, dim3(grid_for(ne, 256)), dim3(256), 0, s, ne, d_tets, d_x, qmin, d_q, first);
  IFEA_HIP(hipGetLastError());
  int h_first = big;
  IFEA_HIP(hipMemcpyAsync(&h_first, first, sizeof(int), hipMemcpyDeviceToHost, s));
  if (q) IFEA_HIP(hipMemcpyAsync(q, d_q, (size_t)ne * sizeof(double), hipMemcpyDeviceToHost, s));
  IFEA_HIP(hipStreamSynchronize(s));
  return h_first == big ? -1 : h_first;
}

void StepDevice::get_v(double* v) {
  IFEA_HIP(hipMemcpyAsync(v, v_.p, ndof_ * sizeof(double), hipMemcpyDeviceToHost, s_));
  IFEA_HIP(hipStreamSynchronize(s_));
}

void StepDevice::set_v(const double* v) {
  IFEA_HIP(hipMemcpyAsync(v_.p, v, ndof_ * sizeof(double), hipMemcpyHostToDevice, s_));
  IFEA_HIP(hipStreamSynchronize(s_));
}

void StepDevice::info(int reset, double* out) {
  for (int k = 0; k < 24; k++) out[k] = 0;
  if (spd_->analyzed()) {
    const Symbolic& S = spd_->symbolic();
    out[0] = (double)S.nnzL;
    out[1] = S.flops;
    out[2] = S.ns;
    out[3] = (double)spd_->device_bytes();
    out[4] = spd_->last_factor_ms();
    out[5] = (double)S.Lsize;
    out[6] = (double)S.Usize;
    out[7] = S.max_level + 1;
  }
  const SpdDevice::Profile& p = spd_->profile();
  out[8] = (double)p.syrk_launches;
  out[9] = p.syrk_ms;
  out[10] = p.syrk_flops;
  out[11] = (double)p.solves;
  out[12] = p.solve_ms;
  out[13] = p.solve_bytes;
  out[14] = (double)p.factorizations;
  out[15] = p.factor_ms;
  out[16] = (double)p.outer_launches;
  out[17] = p.outer_ms;
  out[18] = p.outer_flops;
  out[19] = (double)p.f32_factorizations;
  out[20] = (double)p.f32_solves;
  out[21] = (double)p.outer_launches_f32;
  out[22] = p.outer_ms_f32;
  out[23] = p.outer_flops_f32;
  if (reset) spd_->reset_profile();
}

}  // namespace ifea
