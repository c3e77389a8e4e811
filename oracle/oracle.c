/*
 * oracle.c -- CPU restatement of the reference's level-set reinitialization
 * arithmetic, over the flattened node stencils of casl_reinit.h.
 *
 * TEST INFRASTRUCTURE ONLY.  Only the tests, the smoke check and the
 * benchmark's CPU-baseline leg load this library, and only as the checker;
 * the product path never links or calls it.
 *
 * Built with -O2 -ffp-contract=off (no FMA contraction, no fast-math) so every
 * expression is evaluated in the reference's order with IEEE FP64 rounding.
 * Paths are relative to the reference tree.
 *
 *   MINMOD                               src/casl_math.h:80-88
 *   MIN / MAX (2- and 3-argument)        src/casl_math.h:164-186
 *   interface_Location                   src/casl_math.cpp:55-59
 *   interface_Location_With_Second_Order_Derivative
 *                                        src/casl_math.cpp:67-97
 *   get_linear_interpolation_weights,
 *   interpolate_linearly                 src/my_p4est_quad_neighbor_nodes_of_node.h:222-306
 *   central_second_derivative            src/my_p4est_quad_neighbor_nodes_of_node.h:214-220
 *   laplace                              src/my_p4est_quad_neighbor_nodes_of_node.h:615-635
 *   correct_naive_second_derivatives     src/my_p4est_quad_neighbor_nodes_of_node.cpp:44-298
 *   ngbd_with_quadratic_interpolation    src/my_p4est_quad_neighbor_nodes_of_node.h:1514-1545
 *   f_{m00,...}_linear                   src/my_p4est_quad_neighbor_nodes_of_node.h:1336-1373
 *   second_derivatives_central (1 rank)  src/my_p4est_node_neighbors.cpp:1422-1641
 *   reinitialize_one_iteration           src/my_p4est_level_set.cpp:59-288
 *   reinitialize_within_range_of_phi_0   src/my_p4est_level_set.cpp:290-557
 *   central_derivative                   src/my_p4est_quad_neighbor_nodes_of_node.h:200-206
 *   dd_correction_weight_to_naive_first_derivative
 *                                        src/my_p4est_quad_neighbor_nodes_of_node.h:186-192
 *   naive_d{x,y,z}_needs_*_correction    src/my_p4est_quad_neighbor_nodes_of_node.h:528-568
 *   d{x,y,z}_central_internal            src/my_p4est_quad_neighbor_nodes_of_node.cpp:638-760
 *   inverse_d_max                        src/my_p4est_node_neighbors.cpp:1168-1186
 *   compute_normals                      src/my_p4est_utils.cpp:1720-1740
 *   advect_in_normal_direction(_one_iteration)
 *                                        src/my_p4est_level_set.cpp:559-726
 *   its CFL time step (CF form)          src/my_p4est_level_set.cpp:744-783
 *   compute_mean_curvature(normals)      src/my_p4est_utils.cpp:1648-1690,
 *                                        src/my_p4est_quad_neighbor_nodes_of_node.h:2167-2171
 *
 * With one MPI rank every node is a "local" node and the ghost scatters are
 * no-ops, so the layer/local split of the reference collapses to one loop.
 *
 * Pinning (tests/test_oracle.py): closed forms (uniform-grid second
 * differences; exactness of the corrected second derivatives and of the
 * quadratic neighbor interpolation for quadratic fields on hanging stencils;
 * analytic roots of interface_Location*), invariants (sign preservation,
 * anchoring, linear signed distance as a fixed point, Neumann mirror walls,
 * second-order convergence of the band error) and known-answer scalars.
 * First derivatives: exact for quadratic fields on hanging stencils and
 * (f_p - f_m)/2h on uniform grids; curvature: h*kappa of circles and spheres
 * converges to h/R and 2h/R (tests/test_oracle_derivatives.py).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/casl_reinit.h"

#define EPS 1e-13

static double MINMOD(double a, double b) {
  if (a * b <= 0) return 0;
  if (fabs(a) < fabs(b)) return a;
  return b;
}
static double MIN2(double a, double b) { return a < b ? a : b; }
static double MIN3(double a, double b, double c) { return a < b ? MIN2(a, c) : MIN2(b, c); }
static double MAX2(double a, double b) { return b < a ? a : b; }

double oracle_minmod(double a, double b) { return MINMOD(a, b); }

double oracle_interface_location(double a, double b, double fa, double fb) {
  return 0.5 * (a + b + (a - b) * (fa + fb) / (fb - fa));
}

double oracle_interface_location_2nd(double a, double b, double fa, double fb, double fxxa, double fxxb) {
  if (fabs(fa) < EPS) return a + EPS;
  if (fabs(fb) < EPS) return b - EPS;
  double fxx = MINMOD(fxxa, fxxb);
  double h = b - a;
  double c2 = 0.5 * fxx;
  double c1 = (fb - fa) / h;
  double c0 = 0.5 * (fb + fa) - h * h / 8 * fxx;
  double x;
  if (fabs(c2) < EPS)
    x = -c0 / c1;
  else {
    if (fb < 0)
      x = (-2 * c0) / (c1 - sqrt(c1 * c1 - 4 * c2 * c0));
    else
      x = (-2 * c0) / (c1 + sqrt(c1 * c1 - 4 * c2 * c0));
  }
  return x + 0.5 * (a + b);
}

/* ------------------------------------------------------------ stencil view */
typedef struct {
  int D, F, C, T;
  int64_t N;
  const int32_t *nbr;
  const double *d, *dtr;
  const uint8_t *wall;
  double zdt;
} view_t;

static view_t make_view(const casl_flat_stencil *s) {
  view_t v;
  v.D = s->dim;
  v.F = 2 * s->dim;
  v.C = 1 << (s->dim - 1);
  v.T = 2 * (s->dim - 1);
  v.N = s->n_nodes;
  v.nbr = s->nbr;
  v.d = s->d;
  v.dtr = s->d_tr;
  v.wall = s->wall;
  v.zdt = s->zero_distance_threshold;
  return v;
}

#define NBR(v, f, c, n) ((v)->nbr[((size_t)(f) * (v)->C + (c)) * (v)->N + (n)])
#define DF(v, f, n) ((v)->d[(size_t)(f) * (v)->N + (n)])
#define DT(v, f, k, n) ((v)->dtr[((size_t)(f) * (v)->T + (k)) * (v)->N + (n)])

/* interpolate_linearly: value of g at the face neighbor f of node n. */
static double face_linear(const view_t *v, int f, int64_t n, const double *g) {
  double d_m0 = DT(v, f, 0, n), d_p0 = DT(v, f, 1, n);
  if (v->D == 2) {
    double w_mm = d_p0;
    double w_pm = d_m0;
    double normalization = (d_m0 + d_p0);
    return (g[NBR(v, f, 1, n)] * w_pm + g[NBR(v, f, 0, n)] * w_mm) / normalization;
  } else {
    double d_0m = DT(v, f, 2, n), d_0p = DT(v, f, 3, n);
    double w_mm = d_p0 * d_0p;
    double w_pm = d_m0 * d_0p;
    double w_mp = d_p0 * d_0m;
    double w_pp = d_m0 * d_0m;
    double normalization = (d_m0 + d_p0) * (d_0m + d_0p);
    return (g[NBR(v, f, 3, n)] * w_pp + g[NBR(v, f, 1, n)] * w_pm + g[NBR(v, f, 2, n)] * w_mp +
            g[NBR(v, f, 0, n)] * w_mm) /
           normalization;
  }
}

static double central_second_derivative(double fp, double f0, double fm, double dp, double dm) {
  return ((fp - f0) / dp + (fm - f0) / dm) * 2. / (dp + dm);
}

static int check_if_zero(double value) { return fabs(value) < EPS; }

/* correct_naive_second_derivatives for one value triple */
static void correct_naive(const view_t *v, int64_t n, double *Dxx, double *Dyy, double *Dzz) {
  const double d_m00 = DF(v, 0, n), d_p00 = DF(v, 1, n), d_0m0 = DF(v, 2, n), d_0p0 = DF(v, 3, n);
  const double m12 = DT(v, 0, 0, n) * DT(v, 0, 1, n) / d_m00 / (d_p00 + d_m00) +
                     DT(v, 1, 0, n) * DT(v, 1, 1, n) / d_p00 / (d_p00 + d_m00);
  const double m21 = DT(v, 2, 0, n) * DT(v, 2, 1, n) / d_0m0 / (d_0p0 + d_0m0) +
                     DT(v, 3, 0, n) * DT(v, 3, 1, n) / d_0p0 / (d_0p0 + d_0m0);
  if (v->D == 2) {
    if (check_if_zero(m12)) {
      if (!check_if_zero(m21)) *Dyy -= m21 * *Dxx;
    } else if (check_if_zero(m21)) {
      *Dxx -= m12 * *Dyy;
    } else {
      const double idet = 1.0 / (1.0 - m12 * m21);
      const double naive_Dxx = *Dxx, naive_Dyy = *Dyy;
      *Dxx = naive_Dxx * idet - naive_Dyy * (m12 * idet);
      *Dyy = naive_Dyy * idet - naive_Dxx * (m21 * idet);
    }
    return;
  }
  const double d_00m = DF(v, 4, n), d_00p = DF(v, 5, n);
  const double m13 = DT(v, 0, 2, n) * DT(v, 0, 3, n) / d_m00 / (d_p00 + d_m00) +
                     DT(v, 1, 2, n) * DT(v, 1, 3, n) / d_p00 / (d_p00 + d_m00);
  const double m23 = DT(v, 2, 2, n) * DT(v, 2, 3, n) / d_0m0 / (d_0p0 + d_0m0) +
                     DT(v, 3, 2, n) * DT(v, 3, 3, n) / d_0p0 / (d_0p0 + d_0m0);
  const double m31 = DT(v, 4, 0, n) * DT(v, 4, 1, n) / d_00m / (d_00p + d_00m) +
                     DT(v, 5, 0, n) * DT(v, 5, 1, n) / d_00p / (d_00p + d_00m);
  const double m32 = DT(v, 4, 2, n) * DT(v, 4, 3, n) / d_00m / (d_00p + d_00m) +
                     DT(v, 5, 2, n) * DT(v, 5, 3, n) / d_00p / (d_00p + d_00m);
  if (check_if_zero(m12) && check_if_zero(m13)) {
    if (check_if_zero(m23)) {
      const int m21_non_zero = !check_if_zero(m21);
      const int m31_non_zero = !check_if_zero(m31);
      const int m32_non_zero = !check_if_zero(m32);
      if (m21_non_zero) *Dyy -= m21 * *Dxx;
      if (m31_non_zero) *Dzz -= m31 * *Dxx;
      if (m32_non_zero) *Dzz -= m32 * *Dyy;
    } else if (check_if_zero(m32)) {
      const int m21_non_zero = !check_if_zero(m21);
      const int m31_non_zero = !check_if_zero(m31);
      if (m31_non_zero) *Dzz -= m31 * *Dxx;
      *Dyy -= m23 * *Dzz;
      if (m21_non_zero) *Dyy -= m21 * *Dxx;
    } else {
      const double idet = 1.0 / (1.0 - m23 * m32);
      const double m21_is_zero = check_if_zero(m21);
      const double m31_is_zero = check_if_zero(m31);
      const double naive_Dyy = *Dyy, naive_Dzz = *Dzz;
      *Dyy = ((m21_is_zero ? naive_Dyy : (naive_Dyy - m21 * *Dxx)) -
              m23 * (m31_is_zero ? naive_Dzz : (naive_Dzz - m31 * *Dxx))) *
             idet;
      *Dzz = ((m31_is_zero ? naive_Dzz : (naive_Dzz - m31 * *Dxx)) -
              m32 * (m21_is_zero ? naive_Dyy : (naive_Dyy - m21 * *Dxx))) *
             idet;
    }
  } else if (check_if_zero(m21) && check_if_zero(m23)) {
    if (check_if_zero(m13)) {
      const int m12_non_zero = !check_if_zero(m12);
      const int m31_non_zero = !check_if_zero(m31);
      const int m32_non_zero = !check_if_zero(m32);
      if (m12_non_zero) *Dxx -= m12 * *Dyy;
      if (m31_non_zero) *Dzz -= m31 * *Dxx;
      if (m32_non_zero) *Dzz -= m32 * *Dyy;
    } else if (check_if_zero(m31)) {
      const int m12_non_zero = !check_if_zero(m12);
      const int m32_non_zero = !check_if_zero(m32);
      if (m32_non_zero) *Dzz -= m32 * *Dyy;
      *Dxx -= m13 * *Dzz;
      if (m12_non_zero) *Dxx -= m12 * *Dyy;
    } else {
      const double idet = 1.0 / (1.0 - m13 * m31);
      const int m12_is_zero = check_if_zero(m12);
      const int m32_is_zero = check_if_zero(m32);
      const double naive_Dxx = *Dxx, naive_Dzz = *Dzz;
      *Dxx = ((m12_is_zero ? naive_Dxx : (naive_Dxx - m12 * *Dyy)) -
              m13 * (m32_is_zero ? naive_Dzz : (naive_Dzz - m32 * *Dyy))) *
             idet;
      *Dzz = ((m32_is_zero ? naive_Dzz : (naive_Dzz - m32 * *Dyy)) -
              m31 * (m12_is_zero ? naive_Dxx : (naive_Dxx - m12 * *Dyy))) *
             idet;
    }
  } else if (check_if_zero(m31) && check_if_zero(m32)) {
    if (check_if_zero(m12)) {
      const int m13_non_zero = !check_if_zero(m13);
      const int m21_non_zero = !check_if_zero(m21);
      const int m23_non_zero = !check_if_zero(m23);
      if (m13_non_zero) *Dxx -= m13 * *Dzz;
      if (m21_non_zero) *Dyy -= m21 * *Dxx;
      if (m23_non_zero) *Dyy -= m23 * *Dzz;
    } else if (check_if_zero(m21)) {
      const int m13_non_zero = !check_if_zero(m13);
      const int m23_non_zero = !check_if_zero(m23);
      if (m23_non_zero) *Dyy -= m23 * *Dzz;
      *Dxx -= m12 * *Dyy;
      if (m13_non_zero) *Dxx -= m13 * *Dzz;
    } else {
      const double idet = 1.0 / (1.0 - m12 * m21);
      const int m13_is_zero = check_if_zero(m13);
      const int m23_is_zero = check_if_zero(m23);
      const double naive_Dxx = *Dxx, naive_Dyy = *Dyy;
      *Dxx = ((m13_is_zero ? naive_Dxx : (naive_Dxx - m13 * *Dzz)) -
              m12 * (m23_is_zero ? naive_Dyy : (naive_Dyy - m23 * *Dzz))) *
             idet;
      *Dyy = ((m23_is_zero ? naive_Dyy : (naive_Dyy - m23 * *Dzz)) -
              m21 * (m13_is_zero ? naive_Dxx : (naive_Dxx - m13 * *Dzz))) *
             idet;
    }
  } else {
    const double idet = 1.0 / (1.0 + m12 * m23 * m31 + m13 * m21 * m32 - m12 * m21 - m13 * m31 - m23 * m32);
    const double naive_Dxx = *Dxx, naive_Dyy = *Dyy, naive_Dzz = *Dzz;
    *Dxx = (naive_Dxx * (idet * (1.0 - m23 * m32)) + naive_Dyy * (idet * (m32 * m13 - m12)) +
            naive_Dzz * (idet * (m12 * m23 - m13)));
    *Dyy = (naive_Dxx * (idet * (m31 * m23 - m21)) + naive_Dyy * (idet * (1.0 - m31 * m13)) +
            naive_Dzz * (idet * (m21 * m13 - m23)));
    *Dzz = (naive_Dxx * (idet * (m21 * m32 - m31)) + naive_Dyy * (idet * (m31 * m12 - m32)) +
            naive_Dzz * (idet * (1.0 - m21 * m12)));
  }
}

/* laplace: face values (linear) and corrected second derivatives at node n.
 * fnb[f] for f = m00, p00, 0m0, 0p0[, 00m, 00p]; dd[a] for a = x, y[, z]. */
static void laplace(const view_t *v, int64_t n, const double *g, double *f000, double *fnb, double *dd) {
  *f000 = g[n];
  /* linearly_interpolated_neighbors evaluates p00, m00, 0p0, 0m0, 00p, 00m;
     the face values are independent so the order does not affect them */
  for (int f = 0; f < v->F; ++f) fnb[f] = face_linear(v, f, n, g);
  dd[0] = central_second_derivative(fnb[1], *f000, fnb[0], DF(v, 1, n), DF(v, 0, n));
  dd[1] = central_second_derivative(fnb[3], *f000, fnb[2], DF(v, 3, n), DF(v, 2, n));
  double dzz = 0.0;
  if (v->D == 3) dzz = central_second_derivative(fnb[5], *f000, fnb[4], DF(v, 5, n), DF(v, 4, n));
  correct_naive(v, n, &dd[0], &dd[1], &dzz);
  if (v->D == 3) dd[2] = dzz;
}

/* ngbd_with_quadratic_interpolation for one field */
static void ngbd_quadratic(const view_t *v, int64_t n, const double *g, double *f000, double *fnb, double *dd) {
  laplace(v, n, g, f000, fnb, dd);
  if (v->D == 2) {
    fnb[0] -= (0.5 * DT(v, 0, 0, n) * DT(v, 0, 1, n) * dd[1]);
    fnb[1] -= (0.5 * DT(v, 1, 0, n) * DT(v, 1, 1, n) * dd[1]);
    fnb[2] -= (0.5 * DT(v, 2, 0, n) * DT(v, 2, 1, n) * dd[0]);
    fnb[3] -= (0.5 * DT(v, 3, 0, n) * DT(v, 3, 1, n) * dd[0]);
  } else {
    fnb[0] -= (0.5 * DT(v, 0, 0, n) * DT(v, 0, 1, n) * dd[1] + 0.5 * DT(v, 0, 2, n) * DT(v, 0, 3, n) * dd[2]);
    fnb[1] -= (0.5 * DT(v, 1, 0, n) * DT(v, 1, 1, n) * dd[1] + 0.5 * DT(v, 1, 2, n) * DT(v, 1, 3, n) * dd[2]);
    fnb[2] -= (0.5 * DT(v, 2, 0, n) * DT(v, 2, 1, n) * dd[0] + 0.5 * DT(v, 2, 2, n) * DT(v, 2, 3, n) * dd[2]);
    fnb[3] -= (0.5 * DT(v, 3, 0, n) * DT(v, 3, 1, n) * dd[0] + 0.5 * DT(v, 3, 2, n) * DT(v, 3, 3, n) * dd[2]);
    fnb[4] -= (0.5 * DT(v, 4, 0, n) * DT(v, 4, 1, n) * dd[0] + 0.5 * DT(v, 4, 2, n) * DT(v, 4, 3, n) * dd[1]);
    fnb[5] -= (0.5 * DT(v, 5, 0, n) * DT(v, 5, 1, n) * dd[0] + 0.5 * DT(v, 5, 2, n) * DT(v, 5, 3, n) * dd[1]);
  }
}

/* ------------------------------------------------------ first derivatives */
static double central_derivative(double fp, double f0, double fm, double dp, double dm) {
  return ((fp - f0) * dm / dp + (f0 - fm) * dp / dm) / (dp + dm);
}

static double dd_correction_weight(double d_m, double d_p, double d_m_m, double d_m_p, double d_p_m, double d_p_p) {
  return 0.5 * (d_p_m * d_p_p * d_m / d_p - d_m_m * d_m_p * d_p / d_m) / (d_m + d_p);
}

/* construct_neighbors' inverse_d_max, with its own list of distances */
static double inverse_d_max(const view_t *v, int64_t n) {
  double m = MAX2(DF(v, 0, n), DF(v, 1, n));
  m = MAX2(MAX2(m, DF(v, 2, n)), DF(v, 3, n));
  if (v->D == 3) m = MAX2(MAX2(m, DF(v, 4, n)), DF(v, 5, n));
  m = MAX2(MAX2(m, DT(v, 0, 0, n)), DT(v, 0, 1, n)); /* d_m00_m0, d_m00_p0 */
  m = MAX2(MAX2(m, DT(v, 1, 0, n)), DT(v, 1, 1, n)); /* d_p00_m0, d_p00_p0 */
  m = MAX2(MAX2(m, DT(v, 2, 0, n)), DT(v, 2, 1, n)); /* d_0m0_m0, d_0m0_p0 */
  m = MAX2(MAX2(m, DT(v, 3, 0, n)), DT(v, 3, 1, n)); /* d_0p0_m0, d_0p0_p0 */
  if (v->D == 3) {
    m = MAX2(MAX2(m, DT(v, 0, 2, n)), DT(v, 0, 3, n)); /* d_m00_0m, d_m00_0p */
    m = MAX2(MAX2(m, DT(v, 1, 2, n)), DT(v, 1, 3, n)); /* d_p00_0m, d_p00_0p */
    m = MAX2(MAX2(m, DT(v, 2, 2, n)), DT(v, 2, 3, n)); /* d_0m0_0m, d_0m0_0p */
    m = MAX2(MAX2(m, DT(v, 3, 2, n)), DT(v, 3, 3, n)); /* d_0p0_0m, d_0p0_0p */
    m = MAX2(MAX2(m, DT(v, 4, 0, n)), DT(v, 4, 1, n)); /* d_00m_m0, d_00m_p0 */
    m = MAX2(MAX2(m, DT(v, 4, 2, n)), DT(v, 5, 3, n)); /* d_00m_0m, d_00p_0p (as listed) */
    m = MAX2(MAX2(m, DT(v, 5, 0, n)), DT(v, 5, 1, n)); /* d_00p_m0, d_00p_p0 */
    m = MAX2(MAX2(m, DT(v, 5, 2, n)), DT(v, 5, 3, n)); /* d_00p_0m, d_00p_0p */
  }
  return 1.0 / m;
}

/* d/dx_a of g at node n: naive central derivative of the linear face values,
 * then for each transverse axis (x: y, z; y: x, z; z: x, y) minus its second
 * derivative times the correction weight where that weight is not zero. */
static double d_central(const view_t *v, int64_t n, int a, const double *g) {
  const int fm = 2 * a, fp = 2 * a + 1;
  double fx = central_derivative(face_linear(v, fp, n, g), g[n], face_linear(v, fm, n, g), DF(v, fp, n), DF(v, fm, n));
  const double idm = inverse_d_max(v, n);
  double w[2];
  int need[2] = {0, 0};
  int any = 0;
  for (int k = 0; k < v->D - 1; ++k) {
    w[k] = dd_correction_weight(DF(v, fm, n), DF(v, fp, n), DT(v, fm, 2 * k, n), DT(v, fm, 2 * k + 1, n),
                                DT(v, fp, 2 * k, n), DT(v, fp, 2 * k + 1, n));
    need[k] = !check_if_zero(w[k] * idm);
    any = any || need[k];
  }
  if (any) {
    double f000, fnb[6], dd[3];
    laplace(v, n, g, &f000, fnb, dd);
    static const int transverse[3][2] = {{1, 2}, {0, 2}, {0, 1}};
    for (int k = 0; k < v->D - 1; ++k)
      if (need[k]) fx -= dd[transverse[a][k]] * w[k];
  }
  return fx;
}

/* ------------------------------------------------------------ public API */
int oracle_first_derivatives(const casl_flat_stencil *s, const double *f, double *fd) {
  view_t v = make_view(s);
  const int64_t N = v.N;
#pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < N; ++n)
    for (int a = 0; a < v.D; ++a) fd[(size_t)a * N + n] = d_central(&v, n, a, f);
  return 0;
}

int oracle_normals(const casl_flat_stencil *s, const double *phi, double *nrm) {
  view_t v = make_view(s);
  const int64_t N = v.N;
#pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < N; ++n) {
    double g[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < v.D; ++a) g[a] = d_central(&v, n, a, phi);
    double sq = g[0] * g[0] + g[1] * g[1];
    if (v.D == 3) sq = sq + g[2] * g[2];
    const double abs = sqrt(sq);
    for (int a = 0; a < v.D; ++a) nrm[(size_t)a * N + n] = abs < EPS ? 0.0 : g[a] / abs;
  }
  return 0;
}

int oracle_mean_curvature(const casl_flat_stencil *s, const double *nrm, double *kappa) {
  view_t v = make_view(s);
  const int64_t N = v.N;
#pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < N; ++n) {
    double k = d_central(&v, n, 0, nrm) + d_central(&v, n, 1, nrm + N);
    if (v.D == 3) k = k + d_central(&v, n, 2, nrm + 2 * N);
    kappa[n] = k;
  }
  return 0;
}

int oracle_second_derivatives(const casl_flat_stencil *s, const double *f, double *fdd) {
  view_t v = make_view(s);
  const int64_t N = v.N;
#pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < N; ++n) {
    double f000, fnb[6], dd[3];
    laplace(&v, n, f, &f000, fnb, dd);
    for (int a = 0; a < v.D; ++a) fdd[(size_t)a * N + n] = dd[a];
  }
  return 0;
}

/* ngbd_with_quadratic_interpolation at every node: out[(f)*N + n] for the
 * 2*dim corrected neighbor values, out[F*N + n] = f_000. */
int oracle_ngbd_quadratic(const casl_flat_stencil *s, const double *f, double *out) {
  view_t v = make_view(s);
  const int64_t N = v.N;
#pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < N; ++n) {
    double f000, fnb[6], dd[3];
    ngbd_quadratic(&v, n, f, &f000, fnb, dd);
    for (int k = 0; k < v.F; ++k) out[(size_t)k * N + n] = fnb[k];
    out[(size_t)v.F * N + n] = f000;
  }
  return 0;
}

/* f_{face}_linear of a field at every node: out[f*N + n] */
int oracle_face_linear(const casl_flat_stencil *s, const double *g, double *out) {
  view_t v = make_view(s);
  const int64_t N = v.N;
#pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < N; ++n)
    for (int f = 0; f < v.F; ++f) out[(size_t)f * N + n] = face_linear(&v, f, n, g);
  return 0;
}

/* reinitialize_one_iteration over all nodes; dd0/ddn are NULL for
 * first-order space (second_order == false). mask[n] == 0 copies phi_n. */
static void one_iteration(const view_t *v, double *phi_np1, const double *phi_0, const double *phi_n,
                          double lim_high, double lim_low, const double *dd0, const double *ddn,
                          const uint8_t *mask) {
  const int second_order = (dd0 != NULL && ddn != NULL);
  const int D = v->D, F = v->F;
  const int64_t N = v->N;
  const double zdt = v->zdt;
#pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < N; ++n) {
    if (mask && mask[n] == 0) {
      phi_np1[n] = phi_n[n];
      continue;
    }
    if (fabs(phi_0[n]) < zdt) {
      phi_np1[n] = 0.0;
      continue;
    }
    if (!(lim_low < phi_0[n] && phi_0[n] < lim_high)) {
      phi_np1[n] = phi_0[n];
      continue;
    }
    double p0_000, p0_f[6], p0_dd_unused[3];
    double pn_000, pn_f[6], pn_dd_unused[3];
    ngbd_quadratic(v, n, phi_0, &p0_000, p0_f, p0_dd_unused);
    ngbd_quadratic(v, n, phi_n, &pn_000, pn_f, pn_dd_unused);
    double s[6];
    for (int f = 0; f < F; ++f) s[f] = DF(v, f, n);

    double pn_dd_000[3] = {NAN, NAN, NAN}, pn_dd_f[6] = {NAN, NAN, NAN, NAN, NAN, NAN};
    if (second_order) {
      for (int a = 0; a < D; ++a) pn_dd_000[a] = ddn[(size_t)a * N + n];
      for (int f = 0; f < F; ++f) pn_dd_f[f] = face_linear(v, f, n, ddn + (size_t)(f >> 1) * N);
    }
    int near = 0;
    for (int f = 0; f < F; ++f) near |= (p0_000 * p0_f[f] < 0.0);
    if (near) {
      double p0_dd_000[3] = {NAN, NAN, NAN}, p0_dd_f[6] = {NAN, NAN, NAN, NAN, NAN, NAN};
      if (second_order) {
        for (int a = 0; a < D; ++a) p0_dd_000[a] = dd0[(size_t)a * N + n];
        for (int f = 0; f < F; ++f) p0_dd_f[f] = face_linear(v, f, n, dd0 + (size_t)(f >> 1) * N);
      }
      for (int f = 0; f < F; ++f) {
        if (!(p0_000 * p0_f[f] < 0.0)) continue;
        const int a = f >> 1;
        if ((f & 1) == 0) { /* m side */
          if (second_order)
            s[f] = -oracle_interface_location_2nd(-s[f], 0.0, p0_f[f], p0_000, p0_dd_f[f], p0_dd_000[a]);
          else
            s[f] = -oracle_interface_location(-s[f], 0.0, p0_f[f], p0_000);
        } else { /* p side */
          if (second_order)
            s[f] = oracle_interface_location_2nd(0.0, s[f], p0_000, p0_f[f], p0_dd_000[a], p0_dd_f[f]);
          else
            s[f] = oracle_interface_location(0.0, s[f], p0_000, p0_f[f]);
        }
        pn_f[f] = 0.0;
      }
      for (int f = 0; f < F; ++f) s[f] = MAX2(s[f], zdt);
    }
    /* Neumann boundary condition on the walls */
    const uint8_t w = v->wall[n];
    for (int a = 0; a < D; ++a) {
      const int fm = 2 * a, fp = 2 * a + 1;
      if (w & (1u << fm)) {
        s[fm] = s[fp];
        pn_f[fm] = pn_f[fp];
        if (second_order) pn_dd_000[a] = pn_dd_f[fm] = pn_dd_f[fp] = 0.0;
      }
      if (w & (1u << fp)) {
        s[fp] = s[fm];
        pn_f[fp] = pn_f[fm];
        if (second_order) pn_dd_000[a] = pn_dd_f[fm] = pn_dd_f[fp] = 0.0;
      }
    }
    double Dp[3], Dm[3];
    for (int a = 0; a < D; ++a) {
      Dp[a] = (pn_f[2 * a + 1] - pn_000) / s[2 * a + 1];
      Dm[a] = (pn_000 - pn_f[2 * a]) / s[2 * a];
    }
    if (second_order) {
      for (int a = 0; a < D; ++a) {
        double ddm = MINMOD(pn_dd_f[2 * a], pn_dd_000[a]);
        Dm[a] += 0.5 * s[2 * a] * ddm;
        double ddp = MINMOD(pn_dd_f[2 * a + 1], pn_dd_000[a]);
        Dp[a] -= 0.5 * s[2 * a + 1] * ddp;
      }
    }
    const double sgn = (p0_000 > 0.0 ? 1.0 : -1.0);
    double dt = MIN2(s[0], s[1]);
    dt = MIN3(dt, s[2], s[3]);
    if (D == 3) dt = MIN3(dt, s[4], s[5]);
    dt /= (double)D;
    for (int a = 0; a < D; ++a) {
      if (sgn * Dp[a] > 0.0) Dp[a] = 0.0;
      if (sgn * Dm[a] < 0.0) Dm[a] = 0.0;
    }
    double sum = MAX2(Dp[0] * Dp[0], Dm[0] * Dm[0]) + MAX2(Dp[1] * Dp[1], Dm[1] * Dm[1]);
    if (D == 3) sum = sum + MAX2(Dp[2] * Dp[2], Dm[2] * Dm[2]);
    double out = pn_000 - dt * sgn * (sqrt(sum) - 1.);
    if (p0_000 * out < 0.0) out *= -1.0;
    phi_np1[n] = out;
  }
}

static double largest_dbl_smaller_than_dbl_max(void) {
  volatile double p = 1.0;
  while (DBL_MAX - p == DBL_MAX) p *= 2.0;
  return DBL_MAX - p;
}

/* reinitialize_within_range_of_phi_0; mask may be NULL.
 * Returns 0, or -1 on invalid orders (reference: std::invalid_argument). */
int oracle_reinit(const casl_flat_stencil *s, double *phi, int order_space, int order_time, int n_iter,
                  double limit_high, double limit_low, const uint8_t *mask) {
  if (order_space == 0 || order_space > 2 || order_time == 0 || order_time > 2) return -1;
  view_t v = make_view(s);
  const int64_t N = v.N;
  const int D = v.D;
  const double ldbl = largest_dbl_smaller_than_dbl_max();
  const double zdt = v.zdt;
  const double lim_low = (limit_low < -ldbl ? -DBL_MAX : limit_low - MAX2(zdt, EPS * fabs(limit_low)));
  const double lim_high = (limit_high > ldbl ? DBL_MAX : limit_high + MAX2(zdt, EPS * fabs(limit_high)));

  double *phi_np1 = (double *)malloc(sizeof(double) * (size_t)N);
  double *phi_np2 = (order_time > 1) ? (double *)malloc(sizeof(double) * (size_t)N) : NULL;
  double *phi_0 = (double *)malloc(sizeof(double) * (size_t)N);
  memcpy(phi_0, phi, sizeof(double) * (size_t)N);
  double *dd0 = NULL, *ddn = NULL;
  if (order_space > 1) {
    dd0 = (double *)malloc(sizeof(double) * (size_t)N * D);
    ddn = (double *)malloc(sizeof(double) * (size_t)N * D);
    oracle_second_derivatives(s, phi, dd0);
  }
  for (int i = 0; i < n_iter; ++i) {
    if (order_space > 1) oracle_second_derivatives(s, phi, ddn);
    one_iteration(&v, phi_np1, phi_0, phi, lim_high, lim_low, dd0, ddn, mask);
    if (order_time > 1) {
      if (order_space > 1) oracle_second_derivatives(s, phi_np1, ddn);
      one_iteration(&v, phi_np2, phi_0, phi_np1, lim_high, lim_low, dd0, ddn, mask);
#pragma omp parallel for schedule(static)
      for (int64_t n = 0; n < N; ++n) phi[n] = .5 * (phi[n] + phi_np2[n]);
    } else {
      memcpy(phi, phi_np1, sizeof(double) * (size_t)N);
    }
  }
  free(phi_np1);
  free(phi_np2);
  free(phi_0);
  free(dd0);
  free(ddn);
  return 0;
}

/* Threads the oracle's OpenMP loops use (1 when built without OpenMP). */
/* ---------------------------------------------------------- advection */
/* advect_in_normal_direction_one_iteration at node n */
static double advect_one(const view_t *v, int64_t n, const double *vn, double dt, const double *dd,
                         const double *phi_n) {
  const int D = v->D;
  double p000, pf[6], dloc[3];
  ngbd_quadratic(v, n, phi_n, &p000, pf, dloc);
  double s[6];
  for (int f = 0; f < 2 * D; ++f) s[f] = DF(v, f, n);
  double xx000[3], xxm[3], xxp[3];
  for (int a = 0; a < D; ++a) {
    const double *g = dd + (size_t)a * v->N;
    xx000[a] = g[n];
    xxm[a] = face_linear(v, 2 * a, n, g);
    xxp[a] = face_linear(v, 2 * a + 1, n, g);
  }
  const uint8_t w = v->wall[n];
  for (int a = 0; a < D; ++a) {
    if (w & (1u << (2 * a))) {
      s[2 * a] = s[2 * a + 1];
      pf[2 * a] = pf[2 * a + 1];
      xx000[a] = xxm[a] = xxp[a] = 0.0;
    }
    if (w & (1u << (2 * a + 1))) {
      s[2 * a + 1] = s[2 * a];
      pf[2 * a + 1] = pf[2 * a];
      xx000[a] = xxm[a] = xxp[a] = 0.0;
    }
  }
  double dp[3] = {0.0, 0.0, 0.0}, dm[3] = {0.0, 0.0, 0.0};
  for (int a = 0; a < D; ++a) {
    dp[a] = (pf[2 * a + 1] - p000) / s[2 * a + 1];
    dm[a] = (p000 - pf[2 * a]) / s[2 * a];
  }
  for (int a = 0; a < D; ++a) {
    xxm[a] = MINMOD(xxm[a], xx000[a]);
    dm[a] += 0.5 * s[2 * a] * xxm[a];
    xxp[a] = MINMOD(xxp[a], xx000[a]);
    dp[a] -= 0.5 * s[2 * a + 1] * xxp[a];
  }
  const double sgn = vn[n] > 0.0 ? +1.0 : -1.0;
  for (int a = 0; a < D; ++a) {
    if (sgn * dp[a] > 0.0) dp[a] = 0.0;
    if (sgn * dm[a] < 0.0) dm[a] = 0.0;
  }
  double sum = (dp[0] * dp[0] + dm[0] * dm[0]) + (dp[1] * dp[1] + dm[1] * dm[1]);
  if (D == 3) sum = sum + (dp[2] * dp[2] + dm[2] * dm[2]);
  return p000 - dt * vn[n] * (sqrt(sum));
}

/* advect_in_normal_direction(phi, dt, vn, vn_np1): RK2, phi in place */
int oracle_advect(const casl_flat_stencil *s, double *phi, const double *vn, const double *vn_np1, double dt) {
  view_t v = make_view(s);
  const int64_t N = v.N;
  double *phi_n = (double *)malloc(sizeof(double) * N);
  double *phi_np2 = (double *)malloc(sizeof(double) * N);
  double *dd = (double *)malloc(sizeof(double) * (size_t)v.D * N);
  if (!phi_n || !phi_np2 || !dd) {
    free(phi_n);
    free(phi_np2);
    free(dd);
    return -1;
  }
  oracle_second_derivatives(s, phi, dd);
  memcpy(phi_n, phi, sizeof(double) * N);
#pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < N; ++n) phi[n] = advect_one(&v, n, vn, dt, dd, phi_n);
  oracle_second_derivatives(s, phi, dd);
  const double *v2 = vn_np1 ? vn_np1 : vn;
#pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < N; ++n) phi_np2[n] = advect_one(&v, n, v2, dt, dd, phi);
  for (int64_t n = 0; n < N; ++n) phi[n] = 0.5 * (phi_n[n] + phi_np2[n]);
  free(phi_n);
  free(phi_np2);
  free(dd);
  return 0;
}

/* the CF form's time step: min over nodes of 0.8 |s_min / vn| */
double oracle_advect_cfl_dt(const casl_flat_stencil *s, const double *vn) {
  view_t v = make_view(s);
  double dt = DBL_MAX;
  for (int64_t n = 0; n < v.N; ++n) {
    double s_min = MIN2(fabs(DF(&v, 0, n)), fabs(DF(&v, 1, n)));
    s_min = MIN2(s_min, MIN2(fabs(DF(&v, 2, n)), fabs(DF(&v, 3, n))));
    if (v.D == 3) s_min = MIN2(s_min, MIN2(fabs(DF(&v, 4, n)), fabs(DF(&v, 5, n))));
    dt = MIN2(dt, 0.8 * fabs(s_min / vn[n]));
  }
  return dt;
}


#ifdef _OPENMP
#include <omp.h>
int oracle_num_threads(void) { return omp_get_max_threads(); }
void oracle_set_num_threads(int n) { omp_set_num_threads(n); }
#else
int oracle_num_threads(void) { return 1; }
void oracle_set_num_threads(int n) { (void)n; }
#endif
