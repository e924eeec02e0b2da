// ref_cornelius_driver.cpp -- C entry point over the reference's own
// cornelius v1.3, compiled from the reference tree (see Makefile.ref).  Test
// infrastructure: the checker of the cornelius restatement.
#include "cornelius.cpp"  // from the reference tree, via -I

extern "C" int ref_cornelius_find(int dim, const double *v, const double *dx, double v0,
                                  double *centroids, double *normals) {
  static Cornelius cor;
  double d[4];
  for (int i = 0; i < dim; ++i) d[i] = dx[i];
  cor.init(dim, v0, d);
  int n = 0;
  if (dim == 3) {
    double c[2][2][2];
    double *rows[2][2];
    double **planes[2];
    for (int a = 0; a < 2; ++a) {
      for (int b = 0; b < 2; ++b) {
        for (int e = 0; e < 2; ++e) c[a][b][e] = v[4 * a + 2 * b + e];
        rows[a][b] = c[a][b];
      }
      planes[a] = rows[a];
    }
    cor.find_surface_3d(planes);
  } else if (dim == 4) {
    double c[2][2][2][2];
    double *r3[2][2][2];
    double **r2[2][2];
    double ***r1[2];
    for (int a = 0; a < 2; ++a) {
      for (int b = 0; b < 2; ++b) {
        for (int e = 0; e < 2; ++e) {
          for (int f = 0; f < 2; ++f) c[a][b][e][f] = v[8 * a + 4 * b + 2 * e + f];
          r3[a][b][e] = c[a][b][e];
        }
        r2[a][b] = r3[a][b];
      }
      r1[a] = r2[a];
    }
    cor.find_surface_4d(r1);
  } else {
    return -1;
  }
  n = cor.get_Nelements();
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < dim; ++j) {
      centroids[dim * i + j] = cor.get_centroid_elem(i, j);
      normals[dim * i + j] = cor.get_normal_elem(i, j);
    }
  return n;
}
