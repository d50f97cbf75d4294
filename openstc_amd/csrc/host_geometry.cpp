// Host-side block geometry: ghost coordinates and the 2nd-order metric /
// Jacobian, computed once per block on the CPU in FP64 with the reference's
// operation order (bit-exact; this file is compiled without FMA contraction).
//   ghost grid   : Mesh::init_ghost_grid (src/Mesh.cpp:789-943)
//   metrics      : Block::compute_jac_metric (src/Mesh.cpp:259-521)
#include <cstdint>
#include <vector>

#include "../../include/ostc.h"
#include "ostc_host.h"

namespace {

struct Grid {
  int mx, my, mz, G;
  int64_t nx, ny;
  double *x, *y, *z;
  int64_t at(int i, int j, int k) const { return ((int64_t)(k + G) * ny + (j + G)) * nx + (i + G); }
};

// p(i,j,k) = 2*p(face) - p(mirror): reflection through the boundary node
void reflect(Grid &g, int i, int j, int k, int fi, int fj, int fk, int mi, int mj, int mk) {
  const int64_t d = g.at(i, j, k), f = g.at(fi, fj, fk), m = g.at(mi, mj, mk);
  g.x[d] = 2 * g.x[f] - g.x[m];
  g.y[d] = 2 * g.y[f] - g.y[m];
  g.z[d] = 2 * g.z[f] - g.z[m];
}

// parallelogram rule: p(d) = p(a) + p(b) - p(c)
void parallelogram(Grid &g, int64_t d, int64_t a, int64_t b, int64_t c) {
  g.x[d] = g.x[a] + g.x[b] - g.x[c];
  g.y[d] = g.y[a] + g.y[b] - g.y[c];
  g.z[d] = g.z[a] + g.z[b] - g.z[c];
}

}  // namespace

namespace ostc_host {

void ghost_faces(int mx, int my, int mz, int ngg, double *x, double *y, double *z) {
  Grid g{mx, my, mz, ngg + 1, mx + 2 * (ngg + 1), my + 2 * (ngg + 1), x, y, z};
  for (int k = 0; k < mz; ++k)
    for (int j = 0; j < my; ++j) {
      for (int i = -1; i >= -ngg - 1; --i) reflect(g, i, j, k, 0, j, k, -i, j, k);
      for (int i = mx; i < mx + ngg + 1; ++i) reflect(g, i, j, k, mx - 1, j, k, 2 * (mx - 1) - i, j, k);
    }
  for (int k = 0; k < mz; ++k)
    for (int i = 0; i < mx; ++i) {
      for (int j = -1; j >= -ngg - 1; --j) reflect(g, i, j, k, i, 0, k, i, -j, k);
      for (int j = my; j < my + ngg + 1; ++j) reflect(g, i, j, k, i, my - 1, k, i, 2 * (my - 1) - j, k);
    }
  if (mz == 1) {
    // a 2-D block: the k ghost layers repeat x, y and step z by one per
    // layer (src/Mesh.cpp:840-856)
    for (int j = 0; j < my; ++j)
      for (int i = 0; i < mx; ++i) {
        const int64_t s = g.at(i, j, 0);
        for (int k = -ngg - 1; k <= ngg + 1; ++k) {
          if (k == 0) continue;
          const int64_t d = g.at(i, j, k);
          g.x[d] = g.x[s];
          g.y[d] = g.y[s];
          g.z[d] = g.z[s] + k;
        }
      }
    return;
  }
  for (int j = 0; j < my; ++j)
    for (int i = 0; i < mx; ++i) {
      for (int k = -1; k >= -ngg - 1; --k) reflect(g, i, j, k, i, j, 0, i, j, -k);
      for (int k = mz; k < mz + ngg + 1; ++k) reflect(g, i, j, k, i, j, mz - 1, i, j, 2 * (mz - 1) - k);
    }
}

void ghost_edges_corners(int mx, int my, int mz, int ngg, double *x, double *y, double *z) {
  Grid g{mx, my, mz, ngg + 1, mx + 2 * (ngg + 1), my + 2 * (ngg + 1), x, y, z};
  // edges along k: [-ngg-1,mx+ngg] x ([-ngg-1,-1] | [my,my+ngg]) x [0,mz-1]
  for (int i = 1; i <= ngg + 1; ++i)
    for (int j = 1; j <= ngg + 1; ++j)
      for (int k = 0; k < mz; ++k) {
        const int i1 = -i, j1 = -j, i2 = mx - 1 + i, j2 = my - 1 + j;
        parallelogram(g, g.at(i1, j1, k), g.at(i1, 0, k), g.at(0, j1, k), g.at(0, 0, k));
        parallelogram(g, g.at(i2, j2, k), g.at(i2, my - 1, k), g.at(mx - 1, j2, k), g.at(mx - 1, my - 1, k));
        parallelogram(g, g.at(i1, j2, k), g.at(i1, my - 1, k), g.at(0, j2, k), g.at(0, my - 1, k));
        parallelogram(g, g.at(i2, j1, k), g.at(i2, 0, k), g.at(mx - 1, j1, k), g.at(mx - 1, 0, k));
      }
  // edges along j
  for (int i = 1; i <= ngg + 1; ++i)
    for (int j = 0; j < my; ++j)
      for (int k = 1; k <= ngg + 1; ++k) {
        const int i1 = -i, k1 = -k, i2 = mx - 1 + i, k2 = mz - 1 + k;
        parallelogram(g, g.at(i1, j, k1), g.at(i1, j, 0), g.at(0, j, k1), g.at(0, j, 0));
        parallelogram(g, g.at(i2, j, k2), g.at(i2, j, mz - 1), g.at(mx - 1, j, k2), g.at(mx - 1, j, mz - 1));
        parallelogram(g, g.at(i1, j, k2), g.at(i1, j, mz - 1), g.at(0, j, k2), g.at(0, j, mz - 1));
        parallelogram(g, g.at(i2, j, k1), g.at(i2, j, 0), g.at(mx - 1, j, k1), g.at(mx - 1, j, 0));
      }
  // edges along i and the corners; the reference assigns x for all four,
  // then y, then z -- the coordinates are independent so the order of the
  // components does not change any value
  for (int i = -ngg - 1; i <= mx + ngg; ++i)
    for (int j = 1; j <= ngg + 1; ++j)
      for (int k = 1; k <= ngg + 1; ++k) {
        const int j1 = -j, k1 = -k, j2 = my - 1 + j, k2 = mz - 1 + k;
        parallelogram(g, g.at(i, j1, k1), g.at(i, j1, 0), g.at(i, 0, k1), g.at(i, 0, 0));
        parallelogram(g, g.at(i, j2, k2), g.at(i, j2, mz - 1), g.at(i, my - 1, k2), g.at(i, my - 1, mz - 1));
        parallelogram(g, g.at(i, j1, k2), g.at(i, j1, mz - 1), g.at(i, 0, k2), g.at(i, 0, mz - 1));
        parallelogram(g, g.at(i, j2, k1), g.at(i, j2, 0), g.at(i, my - 1, k1), g.at(i, my - 1, 0));
      }
}

namespace {

struct Metric {
  int mx, my, mz, G, g;
  const double *x, *y, *z;
  double *jac, *met;
  int64_t plane;
  int64_t cat(int i, int j, int k) const {
    return ((int64_t)(k + G) * (my + 2 * G) + (j + G)) * (mx + 2 * G) + (i + G);
  }
  int64_t out(int i, int j, int k) const {
    return ((int64_t)(k + g) * (my + 2 * g) + (j + g)) * (mx + 2 * g) + (i + g);
  }
  // returns false for a non-positive Jacobian (nothing written)
  bool point(int i, int j, int k) const {
    const int64_t ip = cat(i + 1, j, k), im = cat(i - 1, j, k), jp = cat(i, j + 1, k), jm = cat(i, j - 1, k),
                  kp = cat(i, j, k + 1), km = cat(i, j, k - 1);
    const double dxd1 = (x[ip] - x[im]) * 0.5, dyd1 = (y[ip] - y[im]) * 0.5, dzd1 = (z[ip] - z[im]) * 0.5;
    const double dxd2 = (x[jp] - x[jm]) * 0.5, dyd2 = (y[jp] - y[jm]) * 0.5, dzd2 = (z[jp] - z[jm]) * 0.5;
    const double dxd3 = (x[kp] - x[km]) * 0.5, dyd3 = (y[kp] - y[km]) * 0.5, dzd3 = (z[kp] - z[km]) * 0.5;
    const double J = dxd1 * dyd2 * dzd3 + dxd2 * dyd3 * dzd1 + dxd3 * dyd1 * dzd2 - dxd1 * dyd3 * dzd2 -
                     dxd2 * dyd1 * dzd3 - dxd3 * dyd2 * dzd1;
    if (J <= 0) return false;
    const int64_t o = out(i, j, k);
    jac[o] = J;
    const double m[9] = {(dyd2 * dzd3 - dyd3 * dzd2) / J, (dxd3 * dzd2 - dxd2 * dzd3) / J,
                         (dxd2 * dyd3 - dxd3 * dyd2) / J, (dyd3 * dzd1 - dyd1 * dzd3) / J,
                         (dxd1 * dzd3 - dxd3 * dzd1) / J, (dxd3 * dyd1 - dxd1 * dyd3) / J,
                         (dyd1 * dzd2 - dyd2 * dzd1) / J, (dxd2 * dzd1 - dxd1 * dzd2) / J,
                         (dxd1 * dyd2 - dxd2 * dyd1) / J};
    for (int q = 0; q < 9; ++q) met[q * plane + o] = m[q];
    return true;
  }
  void copy_from(int i, int j, int k, int si, int sj, int sk) const {
    const int64_t d = out(i, j, k), s = out(si, sj, sk);
    jac[d] = jac[s];
    for (int q = 0; q < 9; ++q) met[q * plane + d] = met[q * plane + s];
  }
};

}  // namespace

int jac_metric(int mx, int my, int mz, int ngg, const double *x, const double *y, const double *z, double *jac,
               double *metric, int *n_neg) {
  const int g = ngg;
  Metric M{mx, my, mz, ngg + 1, g, x, y, z, jac, metric,
           (int64_t)(mx + 2 * g) * (my + 2 * g) * (mz + 2 * g)};
  int neg = 0;
  // a 2-D block (mz = 1) has its one k plane in the interior pass
  // (src/Mesh.cpp:263-268)
  const int k_min = mz == 1 ? 0 : 1, k_max = mz == 1 ? 1 : mz - 1;
  for (int i = 1; i < mx - 1; ++i)
    for (int j = 1; j < my - 1; ++j)
      for (int k = k_min; k < k_max; ++k)
        if (!M.point(i, j, k)) return -1;
  // ghost slabs in the reference's order; a bad ghost copies its inner neighbour
  auto slab = [&](int i0, int i1, int di, int j0, int j1, int dj, int k0, int k1, int dk, int si, int sj, int sk,
                  int order) {
    // order 0: loops i,j,k ; the reference nests i outer, then j, then k
    (void)order;
    for (int i = i0; di > 0 ? i < i1 : i >= i1; i += di)
      for (int j = j0; dj > 0 ? j < j1 : j >= j1; j += dj)
        for (int k = k0; dk > 0 ? k < k1 : k >= k1; k += dk)
          if (!M.point(i, j, k)) {
            ++neg;
            M.copy_from(i, j, k, i + si, j + sj, k + sk);
          }
  };
  slab(1, mx - 1, 1, 1, my - 1, 1, 0, -ngg, -1, 0, 0, 1, 0);
  slab(1, mx - 1, 1, 1, my - 1, 1, mz - 1, mz + ngg, 1, 0, 0, -1, 0);
  slab(1, mx - 1, 1, 0, -ngg, -1, -ngg, mz + ngg, 1, 0, 1, 0, 0);
  slab(1, mx - 1, 1, my - 1, my + ngg, 1, -ngg, mz + ngg, 1, 0, -1, 0, 0);
  slab(0, -ngg, -1, -ngg, my + ngg, 1, -ngg, mz + ngg, 1, 1, 0, 0, 0);
  slab(mx - 1, mx + ngg, 1, -ngg, my + ngg, 1, -ngg, mz + ngg, 1, -1, 0, 0, 0);
  if (n_neg) *n_neg = neg;
  return 0;
}

}  // namespace ostc_host

extern "C" int ostc_block_geometry(int mx, int my, int mz, int ngg, const double *x_in, const double *y_in,
                                   const double *z_in, double *x_out, double *y_out, double *z_out, double *jac,
                                   double *metric, int *n_neg_jac) {
  if (mx < 3 || my < 3 || (mz < 3 && mz != 1) || ngg < 1 || !x_in || !y_in || !z_in || !jac || !metric)
    return OSTC_ERR_ARG;
  const int G = ngg + 1;
  const int64_t NX = mx + 2 * G, NY = my + 2 * G, NZ = mz + 2 * G;
  std::vector<double> X(NX * NY * NZ, 0.0), Y(X.size(), 0.0), Z(X.size(), 0.0);
  for (int k = 0; k < mz; ++k)
    for (int j = 0; j < my; ++j)
      for (int i = 0; i < mx; ++i) {
        const int64_t s = ((int64_t)k * my + j) * mx + i;
        const int64_t d = ((int64_t)(k + G) * NY + (j + G)) * NX + (i + G);
        X[d] = x_in[s];
        Y[d] = y_in[s];
        Z[d] = z_in[s];
      }
  ostc_host::ghost_faces(mx, my, mz, ngg, X.data(), Y.data(), Z.data());
  ostc_host::ghost_edges_corners(mx, my, mz, ngg, X.data(), Y.data(), Z.data());
  if (ostc_host::jac_metric(mx, my, mz, ngg, X.data(), Y.data(), Z.data(), jac, metric, n_neg_jac) != 0)
    return OSTC_ERR_ARG;
  if (x_out && y_out && z_out) {
    const int g = ngg;
    const int64_t nx = mx + 2 * g, ny = my + 2 * g;
    for (int k = -g; k < mz + g; ++k)
      for (int j = -g; j < my + g; ++j)
        for (int i = -g; i < mx + g; ++i) {
          const int64_t d = ((int64_t)(k + g) * ny + (j + g)) * nx + (i + g);
          const int64_t s = ((int64_t)(k + G) * NY + (j + G)) * NX + (i + G);
          x_out[d] = X[s];
          y_out[d] = Y[s];
          z_out[d] = Z[s];
        }
  }
  return OSTC_OK;
}

// ---------------------------------------------------------------- multi-block

namespace {

int sgn(int v) { return (v > 0) - (v < 0); }

// copy the neighbour block's coordinates into the ghost layers across one
// inner face (Mesh::init_inner_ghost_grid, src/Mesh.cpp:940-985); arrays are
// (ngg+1)-padded
void inner_ghost(const ostc_inner_face &fc, Grid &B, const Grid &N) {
  int min_[3] = {fc.range_start[0], fc.range_start[1], fc.range_start[2]};
  int max_[3] = {fc.range_end[0], fc.range_end[1], fc.range_end[2]};
  int mint_[3] = {fc.target_start[0], fc.target_start[1], fc.target_start[2]};
  const int shift = fc.direction < 0 ? -1 : 1;
  min_[fc.face] += shift;
  max_[fc.face] += shift;
  mint_[fc.target_face] += fc.target_direction > 0 ? -1 : 1;
  const int di1 = fc.loop_dir[0], dj1 = fc.loop_dir[1], dk1 = fc.loop_dir[2];
  int di[3];
  for (int i1 = min_[0]; di1 * (i1 - max_[0]) != 1; i1 += di1) {
    di[0] = i1 - min_[0];
    for (int j1 = min_[1]; dj1 * (j1 - max_[1]) != 1; j1 += dj1) {
      di[1] = j1 - min_[1];
      for (int k1 = min_[2]; dk1 * (k1 - max_[2]) != 1; k1 += dk1) {
        di[2] = k1 - min_[2];
        const int i2 = mint_[0] + fc.target_loop_dir[0] * di[fc.src_tar[0]];
        const int j2 = mint_[1] + fc.target_loop_dir[1] * di[fc.src_tar[1]];
        const int k2 = mint_[2] + fc.target_loop_dir[2] * di[fc.src_tar[2]];
        const int64_t d = B.at(i1, j1, k1), s = N.at(i2, j2, k2);
        B.x[d] = N.x[s];
        B.y[d] = N.y[s];
        B.z[d] = N.z[s];
      }
    }
  }
}

}  // namespace

namespace {
int finish_inner_face(ostc_inner_face f, int ngg, ostc_inner_face *out);
}  // namespace

// InnerFace::register_boundary (src/Mesh.cpp:44-178): the raw record is the
// reference's inner{rank}.txt pair of lines (1-based ranges; the range with
// start == end is the face; in 3-D a negative pair marks the second
// direction, in 2-D the tangential direction is the other in-plane one and
// k is the single plane 0; the target block id is 1-based).
extern "C" int ostc_register_inner_face(const int raw[13], int ngg, int dim, ostc_inner_face *out) {
  if (!raw || !out || ngg < 1 || (dim != 2 && dim != 3)) return OSTC_ERR_ARG;
  ostc_inner_face f{};
  for (int i = 0; i < 3; ++i) {
    f.range_start[i] = raw[2 * i];
    f.range_end[i] = raw[2 * i + 1];
    f.target_start[i] = raw[6 + 2 * i];
    f.target_end[i] = raw[6 + 2 * i + 1];
  }
  f.target_block = raw[12] - 1;
  f.face = -1;
  f.direction = 1;
  f.target_face = -1;
  f.target_direction = 1;
  for (int i = 0; i < dim; ++i)
    if (f.range_start[i] == f.range_end[i]) {
      f.face = i;
      if (f.range_start[i] == 1) f.direction = -1;
      f.range_start[i] -= 1;
      f.range_end[i] -= 1;
      break;
    }
  for (int i = 0; i < dim; ++i)
    if (f.target_start[i] == f.target_end[i]) {
      f.target_face = i;
      if (f.target_start[i] == 1) f.target_direction = -1;
      f.target_start[i] -= 1;
      f.target_end[i] -= 1;
      break;
    }
  if (f.face < 0 || f.target_face < 0) return OSTC_ERR_ARG;
  f.src_tar[f.target_face] = -(f.direction * f.target_direction) * (f.face + 1);
  if (dim == 2) {
    const int next_face = 1 - f.face, target_next_face = 1 - f.target_face;
    f.range_start[next_face] = (f.range_start[next_face] < 0 ? -f.range_start[next_face] : f.range_start[next_face]) - 1;
    f.range_end[next_face] = (f.range_end[next_face] < 0 ? -f.range_end[next_face] : f.range_end[next_face]) - 1;
    f.target_start[next_face] =
        (f.target_start[next_face] < 0 ? -f.target_start[next_face] : f.target_start[next_face]) - 1;
    f.target_end[next_face] = (f.target_end[next_face] < 0 ? -f.target_end[next_face] : f.target_end[next_face]) - 1;
    f.src_tar[target_next_face] = sgn((f.target_end[target_next_face] - f.target_start[target_next_face]) *
                                      (f.range_end[next_face] - f.range_start[next_face])) *
                                  (next_face + 1);
    f.src_tar[2] = 3;
    f.range_start[2] = f.range_end[2] = 0;
    f.target_start[2] = f.target_end[2] = 0;
    return finish_inner_face(f, ngg, out);
  }
  int next_face = -1, target_next_face = -1;
  for (int i = 0; i < 3; ++i)
    if (f.range_start[i] < 0) {
      next_face = i;
      f.range_start[i] = -f.range_start[i] - 1;
      f.range_end[i] = -f.range_end[i] - 1;
      break;
    }
  for (int i = 0; i < 3; ++i)
    if (f.target_start[i] < 0) {
      target_next_face = i;
      f.target_start[i] = -f.target_start[i] - 1;
      f.target_end[i] = -f.target_end[i] - 1;
      break;
    }
  if (next_face < 0 || target_next_face < 0) return OSTC_ERR_ARG;
  f.src_tar[target_next_face] = sgn((f.target_end[target_next_face] - f.target_start[target_next_face]) *
                                    (f.range_end[next_face] - f.range_start[next_face])) *
                                (next_face + 1);
  next_face = 3 - f.face - next_face;
  target_next_face = 3 - f.target_face - target_next_face;
  f.range_start[next_face] -= 1;
  f.range_end[next_face] -= 1;
  f.target_start[target_next_face] -= 1;
  f.target_end[target_next_face] -= 1;
  f.src_tar[target_next_face] = sgn((f.target_end[target_next_face] - f.target_start[target_next_face]) *
                                    (f.range_end[next_face] - f.range_start[next_face])) *
                                (next_face + 1);
  return finish_inner_face(f, ngg, out);
}

namespace {
// loop ranges on the ghost side (corners not communicated), loop directions
// and point counts (src/Mesh.cpp:117-177)
int finish_inner_face(ostc_inner_face f, int ngg, ostc_inner_face *out) {
  f.range_start[f.face] = f.direction < 0 ? -ngg : f.range_end[f.face];
  f.range_end[f.face] = f.direction < 0 ? 0 : f.range_end[f.face] + ngg;
  const int sgn_tar = sgn(f.src_tar[f.target_face]);
  if (f.target_direction > 0)
    f.target_start[f.target_face] = sgn_tar > 0 ? f.target_end[f.target_face] - ngg : f.target_end[f.target_face];
  else
    f.target_start[f.target_face] = sgn_tar > 0 ? 0 : ngg;
  for (int i = 0; i < 3; ++i) {
    f.loop_dir[i] = f.range_start[i] <= f.range_end[i] ? 1 : -1;
    f.target_loop_dir[i] = sgn(f.src_tar[i]);
    f.src_tar[i] = (f.src_tar[i] < 0 ? -f.src_tar[i] : f.src_tar[i]) - 1;
  }
  for (int i = 0; i < 3; ++i) {
    const int d = f.range_start[i] - f.range_end[i];
    f.n_point[i] = (d < 0 ? -d : d) + 1;
  }
  *out = f;
  return OSTC_OK;
}
}  // namespace

// Geometry of a multi-block mesh: ghost faces by reflection, then the ghost
// layers across inner faces copied from the neighbour blocks, then edges and
// corners by the parallelogram rule, then metrics (src/Mesh.cpp:789-862,
// 945-993, 864-937, 259-521, in that order).
extern "C" int ostc_mesh_geometry(int n_blocks, int ngg, ostc_block_coords *blocks, const ostc_inner_face *inner,
                                  const int *n_inner) {
  if (n_blocks < 1 || ngg < 1 || !blocks || (!inner && n_inner)) return OSTC_ERR_ARG;
  const int G = ngg + 1;
  std::vector<std::vector<double>> X(n_blocks), Y(n_blocks), Z(n_blocks);
  std::vector<Grid> grid(n_blocks);
  for (int b = 0; b < n_blocks; ++b) {
    const ostc_block_coords &c = blocks[b];
    if (c.mx < 3 || c.my < 3 || (c.mz < 3 && c.mz != 1) || !c.x || !c.y || !c.z || !c.jac || !c.metric)
      return OSTC_ERR_ARG;
    const int64_t NX = c.mx + 2 * G, NY = c.my + 2 * G, NZ = c.mz + 2 * G;
    X[b].assign(NX * NY * NZ, 0.0);
    Y[b].assign(X[b].size(), 0.0);
    Z[b].assign(X[b].size(), 0.0);
    grid[b] = Grid{c.mx, c.my, c.mz, G, NX, NY, X[b].data(), Y[b].data(), Z[b].data()};
    for (int k = 0; k < c.mz; ++k)
      for (int j = 0; j < c.my; ++j)
        for (int i = 0; i < c.mx; ++i) {
          const int64_t s = ((int64_t)k * c.my + j) * c.mx + i, d = grid[b].at(i, j, k);
          X[b][d] = c.x[s];
          Y[b][d] = c.y[s];
          Z[b][d] = c.z[s];
        }
    ostc_host::ghost_faces(c.mx, c.my, c.mz, ngg, X[b].data(), Y[b].data(), Z[b].data());
  }
  // the neighbours' points across inner faces come before the edges and
  // corners, whose parallelogram rule then starts from them
  // (src/Mesh.cpp:859-862)
  if (n_inner) {
    int off = 0;
    for (int b = 0; b < n_blocks; ++b) {
      for (int f = 0; f < n_inner[b]; ++f) {
        const ostc_inner_face &fc = inner[off + f];
        if (fc.target_block < 0 || fc.target_block >= n_blocks) return OSTC_ERR_ARG;
        inner_ghost(fc, grid[b], grid[fc.target_block]);
      }
      off += n_inner[b];
    }
  }
  for (int b = 0; b < n_blocks; ++b)
    ostc_host::ghost_edges_corners(blocks[b].mx, blocks[b].my, blocks[b].mz, ngg, X[b].data(), Y[b].data(),
                                   Z[b].data());
  for (int b = 0; b < n_blocks; ++b) {
    ostc_block_coords &c = blocks[b];
    if (ostc_host::jac_metric(c.mx, c.my, c.mz, ngg, X[b].data(), Y[b].data(), Z[b].data(), c.jac, c.metric,
                              &c.n_neg_jac) != 0)
      return OSTC_ERR_ARG;
    if (c.x_out && c.y_out && c.z_out)
      for (int k = -ngg; k < c.mz + ngg; ++k)
        for (int j = -ngg; j < c.my + ngg; ++j)
          for (int i = -ngg; i < c.mx + ngg; ++i) {
            const int64_t d = ((int64_t)(k + ngg) * (c.my + 2 * ngg) + (j + ngg)) * (c.mx + 2 * ngg) + (i + ngg);
            const int64_t s = grid[b].at(i, j, k);
            c.x_out[d] = X[b][s];
            c.y_out[d] = Y[b][s];
            c.z_out[d] = Z[b][s];
          }
  }
  return OSTC_OK;
}
