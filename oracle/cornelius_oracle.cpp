// cornelius_oracle.cpp -- CPU restatement of the cornelius v1.3 surface
// finder (freezeout/cornelius-c++-1.3/cornelius.cpp in the reference tree,
// algorithm of Huovinen & Petersen, arXiv:1206.3371).
//
// TEST INFRASTRUCTURE ONLY: the checker of the HIP freeze-out kernels.  Only
// tests/ and bench.py's CPU leg load it.
//
// Semantics followed (reference line numbers in cornelius.cpp):
//  * a cell has elements iff 0 < #corners(value >= v0) < 2^d (:2201-2214,
//    :2257-2272);
//  * square edges are visited in the order (x2=0), (x1=0), (x1=dx), (x2=dx);
//    a cut is the linear zero of value - v0, a corner exactly at v0 next to a
//    corner below it is moved 1e-9 dx along the edge (:1192-1254);
//  * four cuts (ambiguous square) are paired by the mean corner value, and
//    the outside points follow (:1263-1345); otherwise the outside point is
//    the mean of the corners below v0;
//  * a line's normal is (-dy, dx) in its plane, a polygon's is the sum of its
//    triangles' half cross products, a polyhedron's the sum of its
//    tetrahedra's 4-D normals, each part turned towards its own outside point
//    (:91-106, :273-296, :559-624, :815-987);
//  * polygon centroid: vertex mean for three lines, else the area-weighted
//    mean of the triangles (line, vertex mean); polyhedron centroid: the
//    volume-weighted mean of the tetrahedra (line, polygon centroid, vertex
//    mean) (:478-557, :842-926);
//  * lines are joined into polygons (cube) and polygons into polyhedra
//    (hypercube) only when the cell is ambiguous, by matching end points to
//    1e-10 in the L1 norm (:436-476, :743-813, :1552-1650, :1858-1970).
#include <cmath>
#include <cstring>

namespace {

constexpr int D = 4;

struct Line {
  double p[2][D];  // end points (cell-local coordinates)
  double out[D];   // a point on the outside (value < v0)
  int x1, x2;      // the two free axes
  int flip;        // start/end swapped
  const double *start() const { return p[flip]; }
  const double *end() const { return p[1 - flip]; }
};

struct Polygon {
  int cst;             // constant axis of the cube it lies in
  int x1, x2, x3;      // free axes
  const Line *l[24];
  int n;
  double centroid[D];
  double normal[D];
};

void orient(double *nrm, const double *outdir) {
  double dot = 0;
  for (int i = 0; i < D; ++i) dot += outdir[i] * nrm[i];
  if (dot < 0)
    for (int i = 0; i < D; ++i) nrm[i] = -nrm[i];
}

// Square (2-D face): values v[a][b] at (a dx_x1, b dx_x2); constant axes c[2]
// with values cv[2].  Returns the number of lines written to L.
int square_lines(const double v[2][2], const int c[2], const double cv[2], const double *dx,
                 double v0, Line *L) {
  int above = 0;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) above += v[a][b] >= v0;
  if (above == 0 || above == 4) return 0;
  int x1 = -1, x2 = -1;
  for (int i = 0; i < D; ++i)
    if (i != c[0] && i != c[1]) (x1 < 0 ? x1 : x2) = i;
  const double h1 = dx[x1], h2 = dx[x2];
  double cut[4][2];
  int nc = 0;
  // one edge from corner A to corner B; the cut lies at s(A->B)
  auto edge = [&](double a, double b, double fix, bool along_x1, double h) {
    double s;
    if ((a - v0) * (b - v0) < 0) s = (a - v0) / (a - b) * h;
    else if (a == v0 && b < v0) s = 1e-9 * h;
    else if (b == v0 && a < v0) s = (1.0 - 1e-9) * h;
    else return;
    if (along_x1) {
      cut[nc][0] = s;
      cut[nc][1] = fix;
    } else {
      cut[nc][0] = fix;
      cut[nc][1] = s;
    }
    ++nc;
  };
  edge(v[0][0], v[1][0], 0.0, true, h1);
  edge(v[0][0], v[0][1], 0.0, false, h2);
  edge(v[1][0], v[1][1], h1, false, h2);
  edge(v[0][1], v[1][1], h2, true, h1);
  double out[2][2] = {{0, 0}, {0, 0}};
  if (nc == 4) {
    double mid = 0;
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) mid += 0.25 * v[a][b];
    if ((v[0][0] < v0 && mid < v0) || (v[0][0] > v0 && mid > v0))
      for (int k = 0; k < 2; ++k) {
        double t = cut[1][k];
        cut[1][k] = cut[2][k];
        cut[2][k] = t;
      }
    if (mid - v0 < 0) {
      out[0][0] = out[1][0] = 0.5 * h1;
      out[0][1] = out[1][1] = 0.5 * h2;
    } else if (v[0][0] - v0 < 0) {
      out[0][0] = 0;
      out[0][1] = 0;
      out[1][0] = h1;
      out[1][1] = h2;
    } else {
      out[0][0] = h1;
      out[0][1] = 0;
      out[1][0] = 0;
      out[1][1] = h2;
    }
  } else if (nc == 2) {
    int nout = 0;
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        if (v[a][b] < v0) {
          out[0][0] += a * h1;
          out[0][1] += b * h2;
          ++nout;
        }
    if (nout > 0) {
      out[0][0] /= nout;
      out[0][1] /= nout;
    }
  }
  int nl = 0;
  for (int k = 0; k + 1 < nc; k += 2) {
    Line &l = L[nl++];
    for (int e = 0; e < 2; ++e) {
      l.p[e][x1] = cut[k + e][0];
      l.p[e][x2] = cut[k + e][1];
      l.p[e][c[0]] = cv[0];
      l.p[e][c[1]] = cv[1];
    }
    l.out[x1] = out[k / 2][0];
    l.out[x2] = out[k / 2][1];
    l.out[c[0]] = cv[0];
    l.out[c[1]] = cv[1];
    l.x1 = x1;
    l.x2 = x2;
    l.flip = 0;
  }
  return nl;
}

bool same_point(const double *a, const double *b) {
  double d = 0;
  for (int i = 0; i < D; ++i) d += fabs(a[i] - b[i]);
  return d < 1e-10;
}

void polygon_geometry(Polygon &P) {
  double mean[D] = {0, 0, 0, 0};
  for (int i = 0; i < P.n; ++i)
    for (int j = 0; j < D; ++j) mean[j] += P.l[i]->start()[j] + P.l[i]->end()[j];
  for (int j = 0; j < D; ++j) mean[j] = mean[j] / double(2.0 * P.n);
  if (P.n == 3) {
    for (int j = 0; j < D; ++j) P.centroid[j] = mean[j];
  } else {
    double up[D] = {0, 0, 0, 0}, down = 0;
    for (int i = 0; i < P.n; ++i) {
      const double *p1 = P.l[i]->start(), *p2 = P.l[i]->end();
      double a[D], b[D], cm[D];
      for (int j = 0; j < D; ++j) {
        cm[j] = (p1[j] + p2[j] + mean[j]) / 3.0;
        a[j] = p1[j] - mean[j];
        b[j] = p2[j] - mean[j];
      }
      const int x1 = P.x1, x2 = P.x2, x3 = P.x3;
      double A = 0.5 * sqrt(pow(a[x2] * b[x3] - a[x3] * b[x2], 2.0) +
                            pow(a[x1] * b[x3] - a[x3] * b[x1], 2.0) +
                            pow(a[x2] * b[x1] - a[x1] * b[x2], 2.0));
      for (int j = 0; j < D; ++j) up[j] += cm[j] * A;
      down += A;
    }
    for (int j = 0; j < D; ++j) P.centroid[j] = up[j] / down;
  }
  for (int j = 0; j < D; ++j) P.normal[j] = 0;
  for (int i = 0; i < P.n; ++i) {
    const double *p1 = P.l[i]->start(), *p2 = P.l[i]->end();
    double a[D], b[D], nrm[D], o[D];
    for (int j = 0; j < D; ++j) {
      a[j] = p1[j] - P.centroid[j];
      b[j] = p2[j] - P.centroid[j];
    }
    const int x1 = P.x1, x2 = P.x2, x3 = P.x3;
    nrm[x1] = 0.5 * (a[x2] * b[x3] - a[x3] * b[x2]);
    nrm[x2] = -0.5 * (a[x1] * b[x3] - a[x3] * b[x1]);
    nrm[x3] = 0.5 * (a[x1] * b[x2] - a[x2] * b[x1]);
    nrm[P.cst] = 0;
    for (int j = 0; j < D; ++j) o[j] = P.l[i]->out[j] - P.centroid[j];
    orient(nrm, o);
    for (int j = 0; j < D; ++j) P.normal[j] += nrm[j];
  }
}

// Cube (3-D cell or 3-D face of a hypercube): values v[a][b][c] over its
// free axes x1 < x2 < x3, constant axis cst at value cv.  Lines go to L
// (12 at most), polygons to P; returns the polygon count.
struct CubeOut {
  Line L[12];
  int nl;
  Polygon P[6];
  int np;
  int ambiguous;
};

void cube_polygons(const double v[2][2][2], int cst, double cv, const double *dx, double v0,
                   CubeOut &C) {
  int x[3], k = 0;
  for (int i = 0; i < D; ++i)
    if (i != cst) x[k++] = i;
  C.nl = 0;
  C.np = 0;
  C.ambiguous = 0;
  for (int f = 0; f < 3; ++f) {      // free axis held fixed for this face
    for (int j = 0; j < 2; ++j) {
      double sq[2][2];
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
          sq[a][b] = f == 0 ? v[j][a][b] : (f == 1 ? v[a][j][b] : v[a][b][j]);
      const int c[2] = {cst, x[f]};
      const double cvs[2] = {cv, j * dx[x[f]]};
      Line tmp[2];
      const int n = square_lines(sq, c, cvs, dx, v0, tmp);
      // a square has four cuts exactly when it produced two lines
      if (n == 2) C.ambiguous++;
      for (int i = 0; i < n; ++i) C.L[C.nl++] = tmp[i];
    }
  }
  if (C.nl == 0) return;
  if (C.ambiguous == 0 && C.nl == 6) C.ambiguous++;
  auto init_poly = [&](Polygon &P) {
    P.cst = cst;
    P.x1 = x[0];
    P.x2 = x[1];
    P.x3 = x[2];
    P.n = 0;
  };
  if (C.ambiguous > 0) {
    int used = 0;
    bool free_[12];
    for (int i = 0; i < C.nl; ++i) free_[i] = true;
    do {
      Polygon &P = C.P[C.np];
      init_poly(P);
      // restart after every addition; the reference resets its counter to 0
      // and the loop increment then resumes at 1 (line 0 always opens the
      // first polygon, so nothing is lost)
      for (int i = 0; i < C.nl; ++i) {
        if (!free_[i]) continue;
        Line &l = C.L[i];
        bool add = false;
        if (P.n == 0) {
          add = true;
        } else {
          // joined when its start or end meets the last line's end; the end
          // meeting flips the line (checked last, as Polygon::add_line does)
          const double *e = P.l[P.n - 1]->end();
          const bool s_meets = same_point(l.start(), e), e_meets = same_point(l.end(), e);
          if (s_meets || e_meets) {
            if (e_meets) l.flip = 1 - l.flip;
            add = true;
          }
        }
        if (add) {
          P.l[P.n++] = &l;
          free_[i] = false;
          ++used;
          i = 0;
        }
      }
      ++C.np;
    } while (used < C.nl);
  } else {
    Polygon &P = C.P[0];
    init_poly(P);
    for (int i = 0; i < C.nl; ++i) P.l[P.n++] = &C.L[i];
    C.np = 1;
  }
  for (int i = 0; i < C.np; ++i) polygon_geometry(C.P[i]);
}

void tetra_normal(const double *a, const double *b, const double *c, double *n) {
  double bc01 = b[0] * c[1] - b[1] * c[0];
  double bc02 = b[0] * c[2] - b[2] * c[0];
  double bc03 = b[0] * c[3] - b[3] * c[0];
  double bc12 = b[1] * c[2] - b[2] * c[1];
  double bc13 = b[1] * c[3] - b[3] * c[1];
  double bc23 = b[2] * c[3] - b[3] * c[2];
  n[0] = 1.0 / 6.0 * (a[1] * bc23 - a[2] * bc13 + a[3] * bc12);
  n[1] = -1.0 / 6.0 * (a[0] * bc23 - a[2] * bc03 + a[3] * bc02);
  n[2] = 1.0 / 6.0 * (a[0] * bc13 - a[1] * bc03 + a[3] * bc01);
  n[3] = -1.0 / 6.0 * (a[0] * bc12 - a[1] * bc02 + a[2] * bc01);
}

// L1 distance <= 1e-10 (Polyhedron::lines_equal, cornelius.cpp:786-813)
bool same_point_le(const double *a, const double *b) {
  double d = 0;
  for (int i = 0; i < D; ++i) d += fabs(a[i] - b[i]);
  return d <= 1e-10;
}

// lines connected when the start or the end of l1 is the start of l2
bool lines_touch(const Line *l1, const Line *l2) {
  return same_point_le(l1->start(), l2->start()) || same_point_le(l1->end(), l2->start());
}

struct Polyhedron {
  const Polygon *p[48];
  int n;
  int ntet;
  double centroid[D], normal[D];
};
}  // namespace

static void polyhedron_geometry(Polyhedron &H) {
  double mean[D] = {0, 0, 0, 0};
  for (int i = 0; i < H.n; ++i)
    for (int j = 0; j < H.p[i]->n; ++j)
      for (int k = 0; k < D; ++k) mean[k] += H.p[i]->l[j]->start()[k] + H.p[i]->l[j]->end()[k];
  for (int k = 0; k < D; ++k) mean[k] = mean[k] / double(2.0 * H.ntet);
  double up[D] = {0, 0, 0, 0}, down = 0;
  for (int i = 0; i < H.n; ++i) {
    const double *cent = H.p[i]->centroid;
    for (int j = 0; j < H.p[i]->n; ++j) {
      const double *p1 = H.p[i]->l[j]->start(), *p2 = H.p[i]->l[j]->end();
      double a[D], b[D], c[D], n[D], cm[D];
      for (int k = 0; k < D; ++k) {
        cm[k] = (p1[k] + p2[k] + cent[k] + mean[k]) / 4.0;
        a[k] = p1[k] - mean[k];
        b[k] = p2[k] - mean[k];
        c[k] = cent[k] - mean[k];
      }
      tetra_normal(a, b, c, n);
      double V = 0;
      for (int k = 0; k < D; ++k) V += n[k] * n[k];
      V = sqrt(V);
      for (int k = 0; k < D; ++k) up[k] += cm[k] * V;
      down += V;
    }
  }
  for (int k = 0; k < D; ++k) H.centroid[k] = up[k] / down;
  for (int k = 0; k < D; ++k) H.normal[k] = 0;
  for (int i = 0; i < H.n; ++i) {
    const double *cent = H.p[i]->centroid;
    for (int j = 0; j < H.p[i]->n; ++j) {
      const Line *l = H.p[i]->l[j];
      double a[D], b[D], c[D], n[D], o[D];
      for (int k = 0; k < D; ++k) {
        a[k] = l->start()[k] - H.centroid[k];
        b[k] = l->end()[k] - H.centroid[k];
        c[k] = cent[k] - H.centroid[k];
        o[k] = l->out[k] - H.centroid[k];
      }
      tetra_normal(a, b, c, n);
      orient(n, o);
      for (int k = 0; k < D; ++k) H.normal[k] += n[k];
    }
  }
}

extern "C" {

// Surface elements of one cell.  dim = 3: corners v[8], index (it, ix, iy)
// row-major (iy fastest), dx[3]; dim = 4: v[16] over (it, ix, iy, iz), dx[4].
// Writes up to 10 elements' centroid[dim] and normal[dim]; returns the count.
int cfo_find(int dim, const double *v, const double *dxs, double v0, double *centroids,
             double *normals) {
  int above = 0;
  const int nc = 1 << dim;
  for (int i = 0; i < nc; ++i) above += v[i] >= v0;
  if (above == 0 || above == nc) return 0;
  double dx[D];
  for (int i = 0; i < D; ++i) dx[i] = (i < D - dim) ? 1.0 : dxs[i - (D - dim)];
  if (dim == 3) {
    double c[2][2][2];
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        for (int e = 0; e < 2; ++e) c[a][b][e] = v[4 * a + 2 * b + e];
    static thread_local CubeOut C;
    cube_polygons(c, 0, 0.0, dx, v0, C);
    for (int i = 0; i < C.np; ++i)
      for (int j = 0; j < 3; ++j) {
        centroids[3 * i + j] = C.P[i].centroid[j + 1];
        normals[3 * i + j] = C.P[i].normal[j + 1];
      }
    return C.np;
  }
  if (dim != 4) return -1;
  // the eight 3-D faces: axis i held at j (Hypercube::split_to_cubes)
  static thread_local CubeOut C[8];
  const Polygon *polys[48];
  int np = 0, nlines = 0, amb = 0;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 2; ++j) {
      double c[2][2][2];
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
          for (int e = 0; e < 2; ++e) {
            int idx[4];
            int f[3] = {a, b, e}, m = 0;
            for (int k = 0; k < 4; ++k) idx[k] = (k == i) ? j : f[m++];
            c[a][b][e] = v[8 * idx[0] + 4 * idx[1] + 2 * idx[2] + idx[3]];
          }
      CubeOut &Cu = C[2 * i + j];
      cube_polygons(c, i, j * dx[i], dx, v0, Cu);
      for (int k = 0; k < Cu.np; ++k) polys[np++] = &Cu.P[k];
      nlines += Cu.nl;
      if (Cu.ambiguous) ++amb;
    }
  if (amb == 0) {
    int below = 0;
    for (int i = 0; i < 16; ++i) below += v[i] < v0;
    if (below > 8) below = 16 - below;
    if (nlines == 24 && below == 2) ++amb;
  }
  Polyhedron H[10];
  int nh = 0;
  if (amb > 0) {
    bool free_[48];
    for (int i = 0; i < np; ++i) free_[i] = true;
    int used = 0;
    do {
      Polyhedron &Hh = H[nh];
      Hh.n = 0;
      Hh.ntet = 0;
      for (int i = 0; i < np; ++i) {
        if (!free_[i]) continue;
        bool add = Hh.n == 0;
        for (int q = 0; q < Hh.n && !add; ++q)
          for (int a = 0; a < polys[i]->n && !add; ++a)
            for (int b = 0; b < Hh.p[q]->n && !add; ++b)
              if (lines_touch(polys[i]->l[a], Hh.p[q]->l[b])) add = true;
        if (add) {
          Hh.p[Hh.n++] = polys[i];
          Hh.ntet += polys[i]->n;
          free_[i] = false;
          ++used;
          i = 0;
        }
      }
      ++nh;
    } while (used < np && nh < 10);
  } else {
    H[0].n = 0;
    H[0].ntet = 0;
    for (int i = 0; i < np; ++i) {
      H[0].p[H[0].n++] = polys[i];
      H[0].ntet += polys[i]->n;
    }
    nh = 1;
  }
  for (int i = 0; i < nh; ++i) {
    polyhedron_geometry(H[i]);
    for (int j = 0; j < 4; ++j) {
      centroids[4 * i + j] = H[i].centroid[j];
      normals[4 * i + j] = H[i].normal[j];
    }
  }
  return nh;
}

}  // extern "C"
