/* Geometry restatement: ghost coordinates and 2nd-order metrics.
 * TEST INFRASTRUCTURE ONLY (see ostc_oracle.h). */
#include "ostc_oracle.h"

#include <string.h>

void orc_zone_dims(orc_zone *z) {
  z->nx = z->mx + 2 * z->ngg;
  z->ny = z->my + 2 * z->ngg;
  z->nz = z->mz + 2 * z->ngg;
  z->plane = (long)z->nx * z->ny * z->nz;
}

/* index into an array padded by G on every side */
static inline long gidx(int mx, int my, int G, int i, int j, int k) {
  const long nx = mx + 2 * G, ny = my + 2 * G;
  return ((long)(k + G) * ny + (j + G)) * nx + (i + G);
}

#define X(i, j, k) x[gidx(mx, my, G, (i), (j), (k))]
#define Y(i, j, k) y[gidx(mx, my, G, (i), (j), (k))]
#define Z(i, j, k) z[gidx(mx, my, G, (i), (j), (k))]

/* Mesh::init_ghost_grid, single-block part (src/Mesh.cpp:789-857 faces,
 * :864-937 edges/corners by the parallelogram rule).  A block with mz = 1 is
 * a 2-D block: its k ghost layers repeat x, y and step z by one per layer
 * (src/Mesh.cpp:840-856). */
void orc_ghost_grid(int mx, int my, int mz, int ngg, double *x, double *y, double *z) {
  orc_ghost_faces(mx, my, mz, ngg, x, y, z);
  orc_ghost_edges(mx, my, mz, ngg, x, y, z);
}

/* the face layers of Mesh::init_ghost_grid (src/Mesh.cpp:789-857) */
void orc_ghost_faces(int mx, int my, int mz, int ngg, double *x, double *y, double *z) {
  const int G = ngg + 1;
  for (int k = 0; k < mz; ++k)
    for (int j = 0; j < my; ++j) {
      for (int i = -1; i >= -ngg - 1; --i) {
        X(i, j, k) = 2 * X(0, j, k) - X(-i, j, k);
        Y(i, j, k) = 2 * Y(0, j, k) - Y(-i, j, k);
        Z(i, j, k) = 2 * Z(0, j, k) - Z(-i, j, k);
      }
      for (int i = mx; i < mx + ngg + 1; ++i) {
        X(i, j, k) = 2 * X(mx - 1, j, k) - X(2 * (mx - 1) - i, j, k);
        Y(i, j, k) = 2 * Y(mx - 1, j, k) - Y(2 * (mx - 1) - i, j, k);
        Z(i, j, k) = 2 * Z(mx - 1, j, k) - Z(2 * (mx - 1) - i, j, k);
      }
    }
  for (int k = 0; k < mz; ++k)
    for (int i = 0; i < mx; ++i) {
      for (int j = -1; j >= -ngg - 1; --j) {
        X(i, j, k) = 2 * X(i, 0, k) - X(i, -j, k);
        Y(i, j, k) = 2 * Y(i, 0, k) - Y(i, -j, k);
        Z(i, j, k) = 2 * Z(i, 0, k) - Z(i, -j, k);
      }
      for (int j = my; j < my + ngg + 1; ++j) {
        X(i, j, k) = 2 * X(i, my - 1, k) - X(i, 2 * (my - 1) - j, k);
        Y(i, j, k) = 2 * Y(i, my - 1, k) - Y(i, 2 * (my - 1) - j, k);
        Z(i, j, k) = 2 * Z(i, my - 1, k) - Z(i, 2 * (my - 1) - j, k);
      }
    }
  for (int j = 0; j < my; ++j)
    for (int i = 0; i < mx; ++i) {
      if (mz == 1) {
        for (int k = -1; k >= -ngg - 1; --k) {
          X(i, j, k) = X(i, j, 0);
          Y(i, j, k) = Y(i, j, 0);
          Z(i, j, k) = Z(i, j, 0) + k;
        }
        for (int k = mz; k < mz + ngg + 1; ++k) {
          X(i, j, k) = X(i, j, mz - 1);
          Y(i, j, k) = Y(i, j, mz - 1);
          Z(i, j, k) = Z(i, j, mz - 1) + k - mz + 1;
        }
        continue;
      }
      for (int k = -1; k >= -ngg - 1; --k) {
        X(i, j, k) = 2 * X(i, j, 0) - X(i, j, -k);
        Y(i, j, k) = 2 * Y(i, j, 0) - Y(i, j, -k);
        Z(i, j, k) = 2 * Z(i, j, 0) - Z(i, j, -k);
      }
      for (int k = mz; k < mz + ngg + 1; ++k) {
        X(i, j, k) = 2 * X(i, j, mz - 1) - X(i, j, 2 * (mz - 1) - k);
        Y(i, j, k) = 2 * Y(i, j, mz - 1) - Y(i, j, 2 * (mz - 1) - k);
        Z(i, j, k) = 2 * Z(i, j, mz - 1) - Z(i, j, 2 * (mz - 1) - k);
      }
    }
}

/* edges and corners of Mesh::init_ghost_grid by the parallelogram rule
 * (src/Mesh.cpp:864-937); after the inner-face ghost layers in a mesh */
void orc_ghost_edges(int mx, int my, int mz, int ngg, double *x, double *y, double *z) {
  const int G = ngg + 1;
  for (int i = 1; i <= ngg + 1; ++i)
    for (int j = 1; j <= ngg + 1; ++j)
      for (int k = 0; k < mz; ++k) {
        const int i1 = -i, j1 = -j, i2 = mx - 1 + i, j2 = my - 1 + j;
        X(i1, j1, k) = X(i1, 0, k) + X(0, j1, k) - X(0, 0, k);
        Y(i1, j1, k) = Y(i1, 0, k) + Y(0, j1, k) - Y(0, 0, k);
        Z(i1, j1, k) = Z(i1, 0, k) + Z(0, j1, k) - Z(0, 0, k);
        X(i2, j2, k) = X(i2, my - 1, k) + X(mx - 1, j2, k) - X(mx - 1, my - 1, k);
        Y(i2, j2, k) = Y(i2, my - 1, k) + Y(mx - 1, j2, k) - Y(mx - 1, my - 1, k);
        Z(i2, j2, k) = Z(i2, my - 1, k) + Z(mx - 1, j2, k) - Z(mx - 1, my - 1, k);
        X(i1, j2, k) = X(i1, my - 1, k) + X(0, j2, k) - X(0, my - 1, k);
        Y(i1, j2, k) = Y(i1, my - 1, k) + Y(0, j2, k) - Y(0, my - 1, k);
        Z(i1, j2, k) = Z(i1, my - 1, k) + Z(0, j2, k) - Z(0, my - 1, k);
        X(i2, j1, k) = X(i2, 0, k) + X(mx - 1, j1, k) - X(mx - 1, 0, k);
        Y(i2, j1, k) = Y(i2, 0, k) + Y(mx - 1, j1, k) - Y(mx - 1, 0, k);
        Z(i2, j1, k) = Z(i2, 0, k) + Z(mx - 1, j1, k) - Z(mx - 1, 0, k);
      }
  for (int i = 1; i <= ngg + 1; ++i)
    for (int j = 0; j < my; ++j)
      for (int k = 1; k <= ngg + 1; ++k) {
        const int i1 = -i, k1 = -k, i2 = mx - 1 + i, k2 = mz - 1 + k;
        X(i1, j, k1) = X(i1, j, 0) + X(0, j, k1) - X(0, j, 0);
        Y(i1, j, k1) = Y(i1, j, 0) + Y(0, j, k1) - Y(0, j, 0);
        Z(i1, j, k1) = Z(i1, j, 0) + Z(0, j, k1) - Z(0, j, 0);
        X(i2, j, k2) = X(i2, j, mz - 1) + X(mx - 1, j, k2) - X(mx - 1, j, mz - 1);
        Y(i2, j, k2) = Y(i2, j, mz - 1) + Y(mx - 1, j, k2) - Y(mx - 1, j, mz - 1);
        Z(i2, j, k2) = Z(i2, j, mz - 1) + Z(mx - 1, j, k2) - Z(mx - 1, j, mz - 1);
        X(i1, j, k2) = X(i1, j, mz - 1) + X(0, j, k2) - X(0, j, mz - 1);
        Y(i1, j, k2) = Y(i1, j, mz - 1) + Y(0, j, k2) - Y(0, j, mz - 1);
        Z(i1, j, k2) = Z(i1, j, mz - 1) + Z(0, j, k2) - Z(0, j, mz - 1);
        X(i2, j, k1) = X(i2, j, 0) + X(mx - 1, j, k1) - X(mx - 1, j, 0);
        Y(i2, j, k1) = Y(i2, j, 0) + Y(mx - 1, j, k1) - Y(mx - 1, j, 0);
        Z(i2, j, k1) = Z(i2, j, 0) + Z(mx - 1, j, k1) - Z(mx - 1, j, 0);
      }
  for (int i = -ngg - 1; i <= mx + ngg; ++i)
    for (int j = 1; j <= ngg + 1; ++j)
      for (int k = 1; k <= ngg + 1; ++k) {
        const int j1 = -j, k1 = -k, j2 = my - 1 + j, k2 = mz - 1 + k;
        X(i, j1, k1) = X(i, j1, 0) + X(i, 0, k1) - X(i, 0, 0);
        X(i, j2, k2) = X(i, j2, mz - 1) + X(i, my - 1, k2) - X(i, my - 1, mz - 1);
        X(i, j1, k2) = X(i, j1, mz - 1) + X(i, 0, k2) - X(i, 0, mz - 1);
        X(i, j2, k1) = X(i, j2, 0) + X(i, my - 1, k1) - X(i, my - 1, 0);
        Y(i, j1, k1) = Y(i, j1, 0) + Y(i, 0, k1) - Y(i, 0, 0);
        Y(i, j2, k2) = Y(i, j2, mz - 1) + Y(i, my - 1, k2) - Y(i, my - 1, mz - 1);
        Y(i, j1, k2) = Y(i, j1, mz - 1) + Y(i, 0, k2) - Y(i, 0, mz - 1);
        Y(i, j2, k1) = Y(i, j2, 0) + Y(i, my - 1, k1) - Y(i, my - 1, 0);
        Z(i, j1, k1) = Z(i, j1, 0) + Z(i, 0, k1) - Z(i, 0, 0);
        Z(i, j2, k2) = Z(i, j2, mz - 1) + Z(i, my - 1, k2) - Z(i, my - 1, mz - 1);
        Z(i, j1, k2) = Z(i, j1, mz - 1) + Z(i, 0, k2) - Z(i, 0, mz - 1);
        Z(i, j2, k1) = Z(i, j2, 0) + Z(i, my - 1, k1) - Z(i, my - 1, 0);
      }
}

/* One metric point: src/Mesh.cpp:270-299 (all six passes use the same
 * central-difference formula and operation order). */
static int metric_point(int mx, int my, int G, int g, const double *x, const double *y, const double *z, int i,
                        int j, int k, double *jac, double *metric, long plane) {
  const double dxd1 = (X(i + 1, j, k) - X(i - 1, j, k)) * 0.5, dyd1 = (Y(i + 1, j, k) - Y(i - 1, j, k)) * 0.5,
               dzd1 = (Z(i + 1, j, k) - Z(i - 1, j, k)) * 0.5;
  const double dxd2 = (X(i, j + 1, k) - X(i, j - 1, k)) * 0.5, dyd2 = (Y(i, j + 1, k) - Y(i, j - 1, k)) * 0.5,
               dzd2 = (Z(i, j + 1, k) - Z(i, j - 1, k)) * 0.5;
  const double dxd3 = (X(i, j, k + 1) - X(i, j, k - 1)) * 0.5, dyd3 = (Y(i, j, k + 1) - Y(i, j, k - 1)) * 0.5,
               dzd3 = (Z(i, j, k + 1) - Z(i, j, k - 1)) * 0.5;
  const double J = dxd1 * dyd2 * dzd3 + dxd2 * dyd3 * dzd1 + dxd3 * dyd1 * dzd2 - dxd1 * dyd3 * dzd2 -
                   dxd2 * dyd1 * dzd3 - dxd3 * dyd2 * dzd1;
  if (J <= 0) return 0;
  const long id = gidx(mx, my, g, i, j, k);
  jac[id] = J;
  metric[0 * plane + id] = (dyd2 * dzd3 - dyd3 * dzd2) / J;
  metric[1 * plane + id] = (dxd3 * dzd2 - dxd2 * dzd3) / J;
  metric[2 * plane + id] = (dxd2 * dyd3 - dxd3 * dyd2) / J;
  metric[3 * plane + id] = (dyd3 * dzd1 - dyd1 * dzd3) / J;
  metric[4 * plane + id] = (dxd1 * dzd3 - dxd3 * dzd1) / J;
  metric[5 * plane + id] = (dxd3 * dyd1 - dxd1 * dyd3) / J;
  metric[6 * plane + id] = (dyd1 * dzd2 - dyd2 * dzd1) / J;
  metric[7 * plane + id] = (dxd2 * dzd1 - dxd1 * dzd2) / J;
  metric[8 * plane + id] = (dxd1 * dyd2 - dxd2 * dyd1) / J;
  return 1;
}

static void copy_metric(int mx, int my, int g, double *jac, double *metric, long plane, int i, int j, int k, int si,
                        int sj, int sk) {
  const long d = gidx(mx, my, g, i, j, k), s = gidx(mx, my, g, si, sj, sk);
  jac[d] = jac[s];
  for (int l = 0; l < 9; ++l) metric[l * plane + d] = metric[l * plane + s];
}

/* Block::compute_jac_metric (src/Mesh.cpp:259-521): interior, then the six
 * ghost slabs in the reference's order; a non-positive ghost Jacobian copies
 * the inner neighbour's J and metric.  With mz = 1 the interior pass covers
 * k = 0 (src/Mesh.cpp:263-268). */
int orc_jac_metric(int mx, int my, int mz, int ngg, const double *x, const double *y, const double *z,
                   double *jac, double *metric, int *n_neg) {
  const int G = ngg + 1, g = ngg;
  const long plane = (long)(mx + 2 * g) * (my + 2 * g) * (mz + 2 * g);
  int neg = 0;
  const int k_min = mz == 1 ? 0 : 1, k_max = mz == 1 ? 1 : mz - 1;
  for (int i = 1; i < mx - 1; ++i)
    for (int j = 1; j < my - 1; ++j)
      for (int k = k_min; k < k_max; ++k)
        if (!metric_point(mx, my, G, g, x, y, z, i, j, k, jac, metric, plane)) return -1;
  for (int i = 1; i < mx - 1; ++i)
    for (int j = 1; j < my - 1; ++j)
      for (int k = 0; k >= -ngg; --k)
        if (!metric_point(mx, my, G, g, x, y, z, i, j, k, jac, metric, plane)) {
          ++neg;
          copy_metric(mx, my, g, jac, metric, plane, i, j, k, i, j, k + 1);
        }
  for (int i = 1; i < mx - 1; ++i)
    for (int j = 1; j < my - 1; ++j)
      for (int k = mz - 1; k < mz + ngg; ++k)
        if (!metric_point(mx, my, G, g, x, y, z, i, j, k, jac, metric, plane)) {
          ++neg;
          copy_metric(mx, my, g, jac, metric, plane, i, j, k, i, j, k - 1);
        }
  for (int i = 1; i < mx - 1; ++i)
    for (int j = 0; j >= -ngg; --j)
      for (int k = -ngg; k < mz + ngg; ++k)
        if (!metric_point(mx, my, G, g, x, y, z, i, j, k, jac, metric, plane)) {
          ++neg;
          copy_metric(mx, my, g, jac, metric, plane, i, j, k, i, j + 1, k);
        }
  for (int i = 1; i < mx - 1; ++i)
    for (int j = my - 1; j < my + ngg; ++j)
      for (int k = -ngg; k < mz + ngg; ++k)
        if (!metric_point(mx, my, G, g, x, y, z, i, j, k, jac, metric, plane)) {
          ++neg;
          copy_metric(mx, my, g, jac, metric, plane, i, j, k, i, j - 1, k);
        }
  for (int i = 0; i >= -ngg; --i)
    for (int j = -ngg; j < my + ngg; ++j)
      for (int k = -ngg; k < mz + ngg; ++k)
        if (!metric_point(mx, my, G, g, x, y, z, i, j, k, jac, metric, plane)) {
          ++neg;
          copy_metric(mx, my, g, jac, metric, plane, i, j, k, i + 1, j, k);
        }
  for (int i = mx - 1; i < mx + ngg; ++i)
    for (int j = -ngg; j < my + ngg; ++j)
      for (int k = -ngg; k < mz + ngg; ++k)
        if (!metric_point(mx, my, G, g, x, y, z, i, j, k, jac, metric, plane)) {
          ++neg;
          copy_metric(mx, my, g, jac, metric, plane, i, j, k, i - 1, j, k);
        }
  if (n_neg) *n_neg = neg;
  return 0;
}

/* Block::trim_abundant_ghost_mesh (src/Mesh.cpp:534-548) */
void orc_trim_ghost(int mx, int my, int mz, int ngg, const double *src, double *dst) {
  for (int k = -ngg; k < mz + ngg; ++k)
    for (int j = -ngg; j < my + ngg; ++j)
      for (int i = -ngg; i < mx + ngg; ++i) dst[gidx(mx, my, ngg, i, j, k)] = src[gidx(mx, my, ngg + 1, i, j, k)];
}

/* Mesh::init_inner_ghost_grid for one inner face (src/Mesh.cpp:940-985):
 * ghost coordinates of this block across the face copied from the target
 * block's points; both arrays padded by ngg+1. */
void orc_inner_ghost_grid(const orc_inner_face *fc, int mx, int my, int mz, int ngg, double *x, double *y, double *z,
                          int tmx, int tmy, int tmz, const double *tx, const double *ty, const double *tz) {
  const int G = ngg + 1;
  (void)mz;
  (void)tmz;
  int lo[3], hi[3], t0[3];
  for (int d = 0; d < 3; ++d) {
    lo[d] = fc->range_start[d];
    hi[d] = fc->range_end[d];
    t0[d] = fc->target_start[d];
  }
  if (fc->direction < 0) {
    lo[fc->face] -= 1;
    hi[fc->face] -= 1;
  } else {
    lo[fc->face] += 1;
    hi[fc->face] += 1;
  }
  t0[fc->target_face] += fc->target_direction > 0 ? -1 : 1;
  int off[3];
  for (int i = lo[0];; i += fc->loop_dir[0]) {
    off[0] = i - lo[0];
    for (int j = lo[1];; j += fc->loop_dir[1]) {
      off[1] = j - lo[1];
      for (int k = lo[2];; k += fc->loop_dir[2]) {
        off[2] = k - lo[2];
        const int ti = t0[0] + fc->target_loop_dir[0] * off[fc->src_tar[0]];
        const int tj = t0[1] + fc->target_loop_dir[1] * off[fc->src_tar[1]];
        const int tk = t0[2] + fc->target_loop_dir[2] * off[fc->src_tar[2]];
        const long d = gidx(mx, my, G, i, j, k), s = gidx(tmx, tmy, G, ti, tj, tk);
        x[d] = tx[s];
        y[d] = ty[s];
        z[d] = tz[s];
        if (k == hi[2]) break;
      }
      if (j == hi[1]) break;
    }
    if (i == hi[0]) break;
  }
}
