// gvh_cornelius.h -- freeze-out surface elements of one lattice cell on the
// GPU: one wave (64 lanes) per cell, the cell's lines, polygons and
// polyhedra in LDS.
//
// Behaviour of the cornelius v1.3 surface finder (freezeout/cornelius-c++-1.3/
// cornelius.cpp; Huovinen & Petersen, arXiv:1206.3371):
//  * a cell has elements iff 0 < #corners(value >= v0) < 2^d
//    (cornelius.cpp:2201-2214, 2257-2272);
//  * each square (2-D face) is cut on its edges in the order (x2=0), (x1=0),
//    (x1=dx), (x2=dx) by the linear zero of value - v0; a corner exactly at
//    v0 next to one below is moved 1e-9 dx along the edge; four cuts are
//    paired by the mean corner value (cornelius.cpp:1192-1345);
//  * the lines of a cube are joined into polygons only when the cube is
//    ambiguous (a square with four cuts, or six lines), by end points within
//    1e-10 in the L1 norm, reversing a line whose end meets
//    (cornelius.cpp:436-476, 1552-1650); polygons of a hypercube are joined
//    into polyhedra only when it is ambiguous (cornelius.cpp:743-813,
//    1858-1970);
//  * polygon centroid: vertex mean for three lines, else the area-weighted
//    mean of the triangles (line, vertex mean); polyhedron centroid: the
//    volume-weighted mean of the tetrahedra (line, polygon centroid, vertex
//    mean); normals are sums of triangle / tetrahedron normals each turned
//    away from its outside point (cornelius.cpp:91-106, 478-617, 815-987).
//
// Work split: lanes 0..6*ncubes-1 cut one square each (the 8 cubes of a
// hypercube cut their squares separately, as each cube owns its squares); a
// prefix over the wave places the lines; lane q joins the lines of cube q;
// lane g builds polygon g; lane h builds polyhedron h.  The sequential parts
// keep the summation order above and the arithmetic is free of
// contractions, so results equal a plain IEEE-754 double evaluation term by
// term.
#pragma once
#include <hip/hip_runtime.h>

#pragma clang fp contract(off)

namespace gvh {
namespace corn {

constexpr int D = 4;
constexpr int MAXL = 96;    // 8 cubes x 6 squares x 2 lines
constexpr int MAXCP = 12;   // polygons of one cube (one line each at worst)
constexpr int MAXP = 96;    // polygons of a cell
constexpr int MAXE = 10;    // elements of a cell (cornelius.cpp:1703, 2004)

// A line: start, end, and the free-axis components of a point on the outside
// (its two held-axis components equal the start's; the free axes are in
// Work::oax).
struct LineRec {
  double s[D], e[D], o2[2];
};

struct Work {
  LineRec L[MAXL];
  double pcen[MAXP][D];
  unsigned char oax[MAXL];          // free axes x1 | x2 << 2 of each line
  unsigned char ord[MAXL];          // absolute line indices, polygon by polygon
  unsigned char pstart[MAXP], plen[MAXP], pcube[MAXP];
  unsigned char cpst[8][MAXCP], cplen[8][MAXCP];
  unsigned char hlist[MAXP], hfirst[MAXE], hlen[MAXE];
  short hntet[MAXE];
  unsigned char cfirst[8], cnl[8], camb[8], cnp[8];
  int np, nh, nlines, overflow;
  union {
    double tet[MAXL][D + 1];  // per-tetrahedron terms of one polyhedron (4-D cells)
    double pnrm[MAXCP][D];    // polygon normals (3-D cells)
  };
  double v[16];
  double dx[D];
  double v0;
  int dim;
};

// Component i of a 4-vector held in registers, by selects (a run-time index
// into a private array would put the array in scratch memory).
__device__ __forceinline__ double pick(const double *v, int i) {
  return i == 0 ? v[0] : (i == 1 ? v[1] : (i == 2 ? v[2] : v[3]));
}

// Cube q of a cell: the axis held constant, its value and the three free
// axes x0 < x1 < x2.  A 3-D cell is the cube with axis 0 held at 0.
__device__ __forceinline__ void cube_axes(const Work &W, int q, int &cst, double &cv, int &x0,
                                          int &x1, int &x2) {
  if (W.dim == 4) {
    cst = q >> 1;
    cv = (q & 1) * W.dx[cst];
    x0 = cst == 0 ? 1 : 0;
    x1 = cst <= 1 ? 2 : 1;
    x2 = cst <= 2 ? 3 : 2;
  } else {
    cst = 0;
    cv = 0.0;
    x0 = 1;
    x1 = 2;
    x2 = 3;
  }
}

// Corner (a, b, c) of cube q over its free axes: axis k takes q's side j if
// it is the held axis, else the next of a, b, c.
__device__ __forceinline__ double cube_value(const Work &W, int q, int a, int b, int c) {
  if (W.dim == 3) return W.v[4 * a + 2 * b + c];
  const int cst = q >> 1, j = q & 1;
  const int i0 = cst == 0 ? j : a;
  const int i1 = cst == 1 ? j : (cst == 0 ? a : b);
  const int i2 = cst == 2 ? j : (cst <= 1 ? b : c);
  const int i3 = cst == 3 ? j : c;
  return W.v[8 * i0 + 4 * i1 + 2 * i2 + i3];
}

// Cut points of one square with values sq[a][b] at (a h1, b h2) along its
// free axes x1 < x2, in the coordinates of those axes: up to four, in the
// edge order below, and the outside point of each of its nl lines (half the
// number of cut points).  Every index is fixed at compile time, so the
// result stays in registers.
struct SquareCut {
  double cut[4][2];
  double opt[2][2];
  int nl;
};

__device__ __forceinline__ SquareCut square_cut(const double sq[2][2], double h1, double h2,
                                                double v0) {
  SquareCut r;
  r.nl = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) r.cut[k][0] = r.cut[k][1] = 0.0;
  r.opt[0][0] = r.opt[0][1] = r.opt[1][0] = r.opt[1][1] = 0.0;
  int above = 0;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) above += sq[a][b] >= v0;
  if (above == 0 || above == 4) return r;
  // edges from corner value a to corner value b: along x1 at x2 = 0, along
  // x2 at x1 = 0, along x2 at x1 = h1, along x1 at x2 = h2
  const double ea[4] = {sq[0][0], sq[0][0], sq[1][0], sq[0][1]};
  const double eb[4] = {sq[1][0], sq[0][1], sq[1][1], sq[1][1]};
  const double fix[4] = {0.0, 0.0, h1, h2};
  const double hh[4] = {h1, h2, h2, h1};
  const bool along1[4] = {true, false, false, true};
  int nc = 0;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const double a = ea[e], b = eb[e], h = hh[e];
    double s = 0.0;
    bool cut = true;
    if ((a - v0) * (b - v0) < 0) s = (a - v0) / (a - b) * h;
    else if (a == v0 && b < v0) s = 1e-9 * h;
    else if (b == v0 && a < v0) s = (1.0 - 1e-9) * h;
    else cut = false;
    if (cut) {
      const double p0 = along1[e] ? s : fix[e], p1 = along1[e] ? fix[e] : s;
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (k == nc) {
          r.cut[k][0] = p0;
          r.cut[k][1] = p1;
        }
      ++nc;
    }
  }
  if (nc == 4) {
    double mid = 0;
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) mid += 0.25 * sq[a][b];
    if ((sq[0][0] < v0 && mid < v0) || (sq[0][0] > v0 && mid > v0))
      for (int k = 0; k < 2; ++k) {
        const double t = r.cut[1][k];
        r.cut[1][k] = r.cut[2][k];
        r.cut[2][k] = t;
      }
    if (mid - v0 < 0) {
      r.opt[0][0] = r.opt[1][0] = 0.5 * h1;
      r.opt[0][1] = r.opt[1][1] = 0.5 * h2;
    } else if (sq[0][0] - v0 < 0) {
      r.opt[1][0] = h1;
      r.opt[1][1] = h2;
    } else {
      r.opt[0][0] = h1;
      r.opt[1][1] = h2;
    }
  } else if (nc == 2) {
    int nout = 0;
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        if (sq[a][b] < v0) {
          r.opt[0][0] += a * h1;
          r.opt[0][1] += b * h2;
          ++nout;
        }
    if (nout > 0) {
      r.opt[0][0] /= nout;
      r.opt[0][1] /= nout;
    }
  }
  r.nl = nc >> 1;
  return r;
}

// Line K of a square into W.L[i]: free axes x1, x2 take the cut
// coordinates, constant axes c0, c1 the values cv0, cv1.
template <int K>
__device__ __forceinline__ void store_line(Work &W, int i, const SquareCut &r, int c0, int c1,
                                           double cv0, double cv1, int x1, int x2) {
  LineRec &l = W.L[i];
  l.s[x1] = r.cut[2 * K][0];
  l.s[x2] = r.cut[2 * K][1];
  l.s[c0] = cv0;
  l.s[c1] = cv1;
  l.e[x1] = r.cut[2 * K + 1][0];
  l.e[x2] = r.cut[2 * K + 1][1];
  l.e[c0] = cv0;
  l.e[c1] = cv1;
  l.o2[0] = r.opt[K][0];
  l.o2[1] = r.opt[K][1];
  W.oax[i] = (unsigned char)(x1 | (x2 << 2));
}

// Component k of the outside point of line i.
__device__ __forceinline__ double outside(const Work &W, int i, int k) {
  const int a = W.oax[i];
  const LineRec &l = W.L[i];
  return k == (a & 3) ? l.o2[0] : (k == (a >> 2) ? l.o2[1] : l.s[k]);
}

__device__ __forceinline__ bool same_point_lt(const double *a, const double *b) {
  double d = 0;
  for (int i = 0; i < D; ++i) d += fabs(a[i] - b[i]);
  return d < 1e-10;
}

__device__ __forceinline__ bool same_point_le(const double *a, const double *b) {
  double d = 0;
  for (int i = 0; i < D; ++i) d += fabs(a[i] - b[i]);
  return d <= 1e-10;
}

__device__ __forceinline__ void orient(double *nrm, const double *outdir) {
  double dot = 0;
  for (int i = 0; i < D; ++i) dot += outdir[i] * nrm[i];
  if (dot < 0)
    for (int i = 0; i < D; ++i) nrm[i] = -nrm[i];
}

__device__ __forceinline__ void tetra_normal(const double *a, const double *b, const double *c,
                                             double *n) {
  const double bc01 = b[0] * c[1] - b[1] * c[0];
  const double bc02 = b[0] * c[2] - b[2] * c[0];
  const double bc03 = b[0] * c[3] - b[3] * c[0];
  const double bc12 = b[1] * c[2] - b[2] * c[1];
  const double bc13 = b[1] * c[3] - b[3] * c[1];
  const double bc23 = b[2] * c[3] - b[3] * c[2];
  n[0] = 1.0 / 6.0 * (a[1] * bc23 - a[2] * bc13 + a[3] * bc12);
  n[1] = -1.0 / 6.0 * (a[0] * bc23 - a[2] * bc03 + a[3] * bc02);
  n[2] = 1.0 / 6.0 * (a[0] * bc13 - a[1] * bc03 + a[3] * bc01);
  n[3] = -1.0 / 6.0 * (a[0] * bc12 - a[1] * bc02 + a[2] * bc01);
}

__device__ __forceinline__ int pline_index(const Work &W, int g, int k) {
  return W.ord[W.pstart[g] + k];
}
__device__ __forceinline__ const LineRec &pline(const Work &W, int g, int k) {
  return W.L[pline_index(W, g, k)];
}

// Centroid (and, for 3-D cells, normal) of polygon g.
__device__ void polygon_geometry(Work &W, int g, bool want_normal) {
  const int n = W.plen[g];
  int cst, x1, x2, x3;
  double cv;
  cube_axes(W, W.pcube[g], cst, cv, x1, x2, x3);
  double mean[D] = {0, 0, 0, 0};
  for (int i = 0; i < n; ++i) {
    const LineRec &l = pline(W, g, i);
    for (int j = 0; j < D; ++j) mean[j] += l.s[j] + l.e[j];
  }
  for (int j = 0; j < D; ++j) mean[j] = mean[j] / double(2.0 * n);
  double cen[D];
  if (n == 3) {
    for (int j = 0; j < D; ++j) cen[j] = mean[j];
  } else {
    double up[D] = {0, 0, 0, 0}, down = 0;
    for (int i = 0; i < n; ++i) {
      const LineRec &l = pline(W, g, i);
      double a[D], b[D], cm[D];
      for (int j = 0; j < D; ++j) {
        cm[j] = (l.s[j] + l.e[j] + mean[j]) / 3.0;
        a[j] = l.s[j] - mean[j];
        b[j] = l.e[j] - mean[j];
      }
      const double u1 = pick(a, x2) * pick(b, x3) - pick(a, x3) * pick(b, x2);
      const double u2 = pick(a, x1) * pick(b, x3) - pick(a, x3) * pick(b, x1);
      const double u3 = pick(a, x2) * pick(b, x1) - pick(a, x1) * pick(b, x2);
      const double A = 0.5 * sqrt(u1 * u1 + u2 * u2 + u3 * u3);
      for (int j = 0; j < D; ++j) up[j] += cm[j] * A;
      down += A;
    }
    for (int j = 0; j < D; ++j) cen[j] = up[j] / down;
  }
  for (int j = 0; j < D; ++j) W.pcen[g][j] = cen[j];
  if (!want_normal) return;
  double nsum[D] = {0, 0, 0, 0};
  for (int i = 0; i < n; ++i) {
    const int li = pline_index(W, g, i);
    const LineRec &l = W.L[li];
    double a[D], b[D], nrm[D], o[D];
    for (int j = 0; j < D; ++j) {
      a[j] = l.s[j] - cen[j];
      b[j] = l.e[j] - cen[j];
    }
    const double n1 = 0.5 * (pick(a, x2) * pick(b, x3) - pick(a, x3) * pick(b, x2));
    const double n2 = -0.5 * (pick(a, x1) * pick(b, x3) - pick(a, x3) * pick(b, x1));
    const double n3 = 0.5 * (pick(a, x1) * pick(b, x2) - pick(a, x2) * pick(b, x1));
    for (int j = 0; j < D; ++j) nrm[j] = j == x1 ? n1 : (j == x2 ? n2 : (j == x3 ? n3 : 0.0));
    for (int j = 0; j < D; ++j) o[j] = outside(W, li, j) - cen[j];
    orient(nrm, o);
    for (int j = 0; j < D; ++j) nsum[j] += nrm[j];
  }
  for (int j = 0; j < D; ++j) W.pnrm[g][j] = nsum[j];
}

// Polygon g and line j of tetrahedron t of a polyhedron whose np polygons
// are listed from W.hlist[first]: its lines polygon by polygon.
__device__ __forceinline__ void tet_line(const Work &W, int first, int np, int t, int &g, int &j) {
  g = W.hlist[first];
  j = 0;
  for (int i = 0; i < np; ++i) {
    const int gi = W.hlist[first + i];
    const int n = W.plen[gi];
    if (t < n) {
      g = gi;
      j = t;
      return;
    }
    t -= n;
  }
}

// Centroid and normal of polyhedron h, called by all lanes of the wave.  The
// tetrahedra (one per line of each polygon, at most 96) are spread over the
// lanes, lane l taking tetrahedra l and l + 64; their terms go to W.tet and
// every lane sums the rows in tetrahedron order, the order of the
// reference's serial loops, so the sums round the same way.
__device__ void polyhedron_geometry(Work &W, int h, int lane, double *cen_out, double *nrm_out) {
  const int first = W.hfirst[h], np = W.hlen[h], ntet = W.hntet[h];
  static_assert(MAXL <= 128, "two tetrahedra per lane");
  int g[2] = {0, 0}, j[2] = {0, 0}, li[2] = {0, 0};
  const LineRec *ln[2] = {nullptr, nullptr};
#pragma unroll
  for (int r = 0; r < 2; ++r)
    if (lane + 64 * r < ntet) {
      tet_line(W, first, np, lane + 64 * r, g[r], j[r]);
      li[r] = pline_index(W, g[r], j[r]);
      ln[r] = &W.L[li[r]];
    }
  // mean of the line end points
#pragma unroll
  for (int r = 0; r < 2; ++r)
    if (ln[r])
      for (int k = 0; k < D; ++k) W.tet[lane + 64 * r][k] = ln[r]->s[k] + ln[r]->e[k];
  __syncthreads();
  double mean[D] = {0, 0, 0, 0};
#pragma unroll 4
  for (int t = 0; t < ntet; ++t)
    for (int k = 0; k < D; ++k) mean[k] += W.tet[t][k];
  for (int k = 0; k < D; ++k) mean[k] = mean[k] / double(2.0 * ntet);
  __syncthreads();
  // volume-weighted centroid
#pragma unroll
  for (int r = 0; r < 2; ++r)
    if (ln[r]) {
      const LineRec &l = *ln[r];
      const double *pc = W.pcen[g[r]];
      double a[D], b[D], c[D], n[D], cm[D];
      for (int k = 0; k < D; ++k) {
        cm[k] = (l.s[k] + l.e[k] + pc[k] + mean[k]) / 4.0;
        a[k] = l.s[k] - mean[k];
        b[k] = l.e[k] - mean[k];
        c[k] = pc[k] - mean[k];
      }
      tetra_normal(a, b, c, n);
      double V = 0;
      for (int k = 0; k < D; ++k) V += n[k] * n[k];
      V = sqrt(V);
      for (int k = 0; k < D; ++k) W.tet[lane + 64 * r][k] = cm[k] * V;
      W.tet[lane + 64 * r][D] = V;
    }
  __syncthreads();
  double up[D] = {0, 0, 0, 0}, down = 0;
#pragma unroll 4
  for (int t = 0; t < ntet; ++t) {
    for (int k = 0; k < D; ++k) up[k] += W.tet[t][k];
    down += W.tet[t][D];
  }
  double cen[D];
  for (int k = 0; k < D; ++k) cen[k] = up[k] / down;
  __syncthreads();
  // normal: the oriented tetrahedron normals about the centroid
#pragma unroll
  for (int r = 0; r < 2; ++r)
    if (ln[r]) {
      const LineRec &l = *ln[r];
      const double *pc = W.pcen[g[r]];
      double a[D], b[D], c[D], n[D], o[D];
      for (int k = 0; k < D; ++k) {
        a[k] = l.s[k] - cen[k];
        b[k] = l.e[k] - cen[k];
        c[k] = pc[k] - cen[k];
        o[k] = outside(W, li[r], k) - cen[k];
      }
      tetra_normal(a, b, c, n);
      orient(n, o);
      for (int k = 0; k < D; ++k) W.tet[lane + 64 * r][k] = n[k];
    }
  __syncthreads();
  double nrm[D] = {0, 0, 0, 0};
#pragma unroll 4
  for (int t = 0; t < ntet; ++t)
    for (int k = 0; k < D; ++k) nrm[k] += W.tet[t][k];
  if (lane == 0)
    for (int k = 0; k < D; ++k) {
      cen_out[k] = cen[k];
      nrm_out[k] = nrm[k];
    }
  __syncthreads();
}

// Lines joined into polyhedra when the start or the end of a line of the
// candidate polygon is the start of a line already in the polyhedron.
__device__ bool polygons_touch(const Work &W, int gi, int gq) {
  for (int a = 0; a < W.plen[gi]; ++a) {
    const LineRec &l1 = pline(W, gi, a);
    for (int b = 0; b < W.plen[gq]; ++b) {
      const LineRec &l2 = pline(W, gq, b);
      if (same_point_le(l1.s, l2.s) || same_point_le(l1.e, l2.s)) return true;
    }
  }
  return false;
}

// Elements of the cell whose corners W.v, spacings W.dx, iso-value W.v0 and
// dimension W.dim are set.  Called by all 64 lanes of a one-wave workgroup;
// returns the element count (<= MAXE) to every lane and leaves centroid i /
// normal i in cen[i], nrm[i] (4-vectors; a 3-D cell uses axes 1..3).
__device__ int cell_elements(Work &W, double (*cen)[D], double (*nrm)[D]) {
  const int lane = threadIdx.x;
  const int ncube = W.dim == 4 ? 8 : 1;
  // no element unless the cell has corners on both sides
  {
    const int nc = 1 << W.dim;
    int above = 0;
    for (int i = 0; i < nc; ++i) above += W.v[i] >= W.v0;
    if (above == 0 || above == nc) return 0;
  }
  // --- 1: squares -> lines
  SquareCut sc;
  sc.nl = 0;
  int c0 = 0, c1 = 0, x1 = 0, x2 = 0;
  double cv0 = 0.0, cv1 = 0.0;
  const int q = lane / 6, r = lane % 6, f = r >> 1, jj = r & 1;
  if (lane < 6 * ncube) {
    // square (f, jj) of cube q: free axis f held at side jj
    int cst, a0, a1, a2;
    double cv;
    cube_axes(W, q, cst, cv, a0, a1, a2);
    double sq[2][2];
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        sq[a][b] = f == 0 ? cube_value(W, q, jj, a, b)
                          : (f == 1 ? cube_value(W, q, a, jj, b) : cube_value(W, q, a, b, jj));
    const int xf = f == 0 ? a0 : (f == 1 ? a1 : a2);
    x1 = f == 0 ? a1 : a0;
    x2 = f == 2 ? a1 : a2;
    c0 = cst;
    c1 = xf;
    cv0 = cv;
    cv1 = jj * W.dx[xf];
    sc = square_cut(sq, W.dx[x1], W.dx[x2], W.v0);
  }
  const int cnt = sc.nl;
  const unsigned long long b1 = __ballot(cnt >= 1), b2 = __ballot(cnt == 2);
  const unsigned long long lt = (1ull << lane) - 1ull;
  const int pos = __popcll(b1 & lt) + __popcll(b2 & lt);
  if (cnt >= 1) store_line<0>(W, pos, sc, c0, c1, cv0, cv1, x1, x2);
  if (cnt == 2) store_line<1>(W, pos + 1, sc, c0, c1, cv0, cv1, x1, x2);
  if (lane < ncube) {
    const unsigned long long below = (1ull << (6 * lane)) - 1ull;
    const unsigned long long mine = 0x3full << (6 * lane);
    W.cfirst[lane] = __popcll(b1 & below) + __popcll(b2 & below);
    W.cnl[lane] = __popcll(b1 & mine) + __popcll(b2 & mine);
    W.camb[lane] = __popcll(b2 & mine);
  }
  if (lane == 0) W.nlines = __popcll(b1) + __popcll(b2);
  __syncthreads();
  // --- 2: lines of each cube -> polygons
  if (lane < ncube) {
    const int nl = W.cnl[lane], first = W.cfirst[lane];
    int amb = W.camb[lane], np = 0;
    if (nl > 0) {
      if (amb == 0 && nl == 6) amb = 1;
      if (amb > 0) {
        unsigned freem = (1u << nl) - 1u;
        int used = 0, pos2 = 0;
        do {
          const int start = pos2;
          int n = 0;
          double le[D] = {0, 0, 0, 0};
          // after an addition the scan restarts; the reference resets its
          // counter to 0 and the loop increment resumes it at 1
          for (int i = 0; i < nl; ++i) {
            if (!((freem >> i) & 1u)) continue;
            LineRec &l = W.L[first + i];
            bool add = n == 0;
            if (!add) {
              const bool sm = same_point_lt(l.s, le), em = same_point_lt(l.e, le);
              if (sm || em) {
                if (em)
                  for (int k = 0; k < D; ++k) {
                    const double t = l.s[k];
                    l.s[k] = l.e[k];
                    l.e[k] = t;
                  }
                add = true;
              }
            }
            if (add) {
              W.ord[first + pos2++] = (unsigned char)(first + i);
              freem &= ~(1u << i);
              ++used;
              ++n;
              for (int k = 0; k < D; ++k) le[k] = l.e[k];
              i = 0;
            }
          }
          W.cpst[lane][np] = (unsigned char)(first + start);
          W.cplen[lane][np] = (unsigned char)n;
          ++np;
        } while (used < nl);
      } else {
        for (int i = 0; i < nl; ++i) W.ord[first + i] = (unsigned char)(first + i);
        W.cpst[lane][0] = (unsigned char)first;
        W.cplen[lane][0] = (unsigned char)nl;
        np = 1;
      }
    }
    W.camb[lane] = amb;
    W.cnp[lane] = np;
  }
  __syncthreads();
  if (lane == 0) {
    int g = 0;
    for (int c = 0; c < ncube; ++c)
      for (int p = 0; p < W.cnp[c]; ++p, ++g) {
        W.pstart[g] = W.cpst[c][p];
        W.plen[g] = W.cplen[c][p];
        W.pcube[g] = (unsigned char)c;
      }
    W.np = g;
  }
  __syncthreads();
  // --- 3: polygon centroids (and normals of a 3-D cell's polygons)
  for (int g = lane; g < W.np; g += 64) polygon_geometry(W, g, W.dim == 3);
  __syncthreads();
  if (W.dim == 3) {
    const int ne = W.np < MAXE ? W.np : MAXE;
    if (lane < ne)
      for (int k = 0; k < D; ++k) {
        cen[lane][k] = W.pcen[lane][k];
        nrm[lane][k] = W.pnrm[lane][k];
      }
    if (lane == 0 && W.np > MAXE) W.overflow = 1;
    return ne;
  }
  // --- 4: polygons -> polyhedra
  if (lane == 0) {
    int amb = 0;
    for (int c = 0; c < 8; ++c)
      if (W.camb[c]) ++amb;
    if (amb == 0) {
      int below = 0;
      for (int i = 0; i < 16; ++i) below += W.v[i] < W.v0;
      if (below > 8) below = 16 - below;
      if (W.nlines == 24 && below == 2) ++amb;
    }
    const int np = W.np;
    int nh = 0;
    if (amb > 0) {
      unsigned long long fr0 = ~0ull, fr1 = ~0ull;  // free polygons (up to 96)
      auto is_free = [&](int i) { return i < 64 ? ((fr0 >> i) & 1ull) : ((fr1 >> (i - 64)) & 1ull); };
      auto take = [&](int i) {
        if (i < 64) fr0 &= ~(1ull << i);
        else fr1 &= ~(1ull << (i - 64));
      };
      int used = 0, pos3 = 0;
      do {
        const int first = pos3;
        int nin = 0, ntet = 0;
        for (int i = 0; i < np; ++i) {
          if (!is_free(i)) continue;
          bool add = nin == 0;
          for (int k = 0; k < nin && !add; ++k)
            if (polygons_touch(W, i, W.hlist[first + k])) add = true;
          if (add) {
            W.hlist[pos3++] = (unsigned char)i;
            ++nin;
            ntet += W.plen[i];
            take(i);
            ++used;
            i = 0;
          }
        }
        W.hfirst[nh] = (unsigned char)first;
        W.hlen[nh] = (unsigned char)nin;
        W.hntet[nh] = (short)ntet;
        ++nh;
      } while (used < np && nh < MAXE);
      if (used < np) W.overflow = 1;
    } else {
      int ntet = 0;
      for (int i = 0; i < np; ++i) {
        W.hlist[i] = (unsigned char)i;
        ntet += W.plen[i];
      }
      W.hfirst[0] = 0;
      W.hlen[0] = (unsigned char)np;
      W.hntet[0] = (short)ntet;
      nh = 1;
    }
    W.nh = nh;
  }
  __syncthreads();
  // --- 5: polyhedron centroids and normals
  const int nh = W.nh;
  for (int h = 0; h < nh; ++h) polyhedron_geometry(W, h, lane, cen[h], nrm[h]);
  return nh;
}

}  // namespace corn
}  // namespace gvh
