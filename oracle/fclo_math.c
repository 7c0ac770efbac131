/* fclo_math.c — small fixed-size linear algebra and the Jacobi eigen-solver
 * of the oracle (test infrastructure only; see fclo.h for the fixed
 * evaluation order). */
#include <math.h>
#include <stdio.h>
#include "fclo.h"

double fclo_dot(const double a[3], const double b[3]) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

void fclo_cross(const double a[3], const double b[3], double out[3]) {
  double x = a[1] * b[2] - a[2] * b[1];
  double y = a[2] * b[0] - a[0] * b[2];
  double z = a[0] * b[1] - a[1] * b[0];
  out[0] = x;
  out[1] = y;
  out[2] = z;
}

/* math/geometry-inl.h:475-558 (eigen_old): cyclic Jacobi, at most 50 sweeps.
 * vout receives the eigenvectors in the reference's layout: vout.col(k) holds
 * row k of the Jacobi accumulator v, i.e. vout(r,k) = v[k][r]. */
void fclo_eigen_old(const double m[9], double dout[3], double vout[9]) {
  double R[3][3];
  int n = 3, j, iq, ip, i;
  double tresh, theta, tau, t, sm, s, h, g, c;
  double b[3], z[3], d[3];
  double v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (i = 0; i < 3; ++i)
    for (j = 0; j < 3; ++j) R[i][j] = m[3 * i + j];

  for (ip = 0; ip < n; ++ip) {
    b[ip] = d[ip] = R[ip][ip];
    z[ip] = 0;
  }

  for (i = 0; i < 50; ++i) {
    sm = 0;
    for (ip = 0; ip < n; ++ip)
      for (iq = ip + 1; iq < n; ++iq) sm += fabs(R[ip][iq]);
    if (sm == 0.0) {
      for (int k = 0; k < 3; ++k)
        for (int r = 0; r < 3; ++r) vout[3 * r + k] = v[k][r];
      dout[0] = d[0];
      dout[1] = d[1];
      dout[2] = d[2];
      return;
    }

    if (i < 3)
      tresh = 0.2 * sm / (n * n);
    else
      tresh = 0.0;

    for (ip = 0; ip < n; ++ip) {
      for (iq = ip + 1; iq < n; ++iq) {
        g = 100.0 * fabs(R[ip][iq]);
        if (i > 3 && fabs(d[ip]) + g == fabs(d[ip]) &&
            fabs(d[iq]) + g == fabs(d[iq])) {
          R[ip][iq] = 0.0;
        } else if (fabs(R[ip][iq]) > tresh) {
          h = d[iq] - d[ip];
          if (fabs(h) + g == fabs(h)) {
            t = (R[ip][iq]) / h;
          } else {
            theta = 0.5 * h / (R[ip][iq]);
            t = 1.0 / (fabs(theta) + sqrt(1.0 + theta * theta));
            if (theta < 0.0) t = -t;
          }
          c = 1.0 / sqrt(1 + t * t);
          s = t * c;
          tau = s / (1.0 + c);
          h = t * R[ip][iq];
          z[ip] -= h;
          z[iq] += h;
          d[ip] -= h;
          d[iq] += h;
          R[ip][iq] = 0.0;
          /* the three off-diagonal sweeps and the accumulator update */
          for (j = 0; j < ip; ++j) {
            g = R[j][ip]; h = R[j][iq];
            R[j][ip] = g - s * (h + g * tau);
            R[j][iq] = h + s * (g - h * tau);
          }
          for (j = ip + 1; j < iq; ++j) {
            g = R[ip][j]; h = R[j][iq];
            R[ip][j] = g - s * (h + g * tau);
            R[j][iq] = h + s * (g - h * tau);
          }
          for (j = iq + 1; j < n; ++j) {
            g = R[ip][j]; h = R[iq][j];
            R[ip][j] = g - s * (h + g * tau);
            R[iq][j] = h + s * (g - h * tau);
          }
          for (j = 0; j < n; ++j) {
            g = v[j][ip]; h = v[j][iq];
            v[j][ip] = g - s * (h + g * tau);
            v[j][iq] = h + s * (g - h * tau);
          }
        }
      }
    }
    for (ip = 0; ip < n; ++ip) {
      b[ip] += z[ip];
      d[ip] = b[ip];
      z[ip] = 0.0;
    }
  }
  /* The reference prints this and returns with its outputs unset; the
   * oracle returns the current accumulator so that results stay defined. */
  fprintf(stderr, "eigen: too many iterations in Jacobi transform.\n");
  for (int k = 0; k < 3; ++k)
    for (int r = 0; r < 3; ++r) vout[3 * r + k] = v[k][r];
  dout[0] = d[0];
  dout[1] = d[1];
  dout[2] = d[2];
}

/* math/geometry-inl.h:653-682: R = R1^T R2, T = R1^T (t2 - t1). */
void fclo_relative_transform(const double tf1[12], const double tf2[12],
                             double R[9], double T[3]) {
  const double *R1 = tf1, *R2 = tf2;
  double d[3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      R[3 * i + j] = (R1[0 + i] * R2[0 + j] + R1[3 + i] * R2[3 + j]) +
                     R1[6 + i] * R2[6 + j];
  for (int i = 0; i < 3; ++i) d[i] = tf2[9 + i] - tf1[9 + i];
  for (int i = 0; i < 3; ++i)
    T[i] = (R1[0 + i] * d[0] + R1[3 + i] * d[1]) + R1[6 + i] * d[2];
}

/* mesh_distance_traversal_node-inl.h:632: tf = tf1.inverse(Isometry) * tf2.
 * The isometry inverse is (R1^T, -(R1^T t1)); the product is
 * (R1^T R2, R1^T t2 + (-(R1^T t1))). */
void fclo_inverse_compose(const double tf1[12], const double tf2[12],
                          double out[12]) {
  const double *R1 = tf1, *R2 = tf2;
  double it[3], a[3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      out[3 * i + j] = (R1[0 + i] * R2[0 + j] + R1[3 + i] * R2[3 + j]) +
                       R1[6 + i] * R2[6 + j];
  for (int i = 0; i < 3; ++i) {
    it[i] = -((R1[0 + i] * tf1[9] + R1[3 + i] * tf1[10]) + R1[6 + i] * tf1[11]);
    a[i] = (R1[0 + i] * tf2[9] + R1[3 + i] * tf2[10]) + R1[6 + i] * tf2[11];
  }
  for (int i = 0; i < 3; ++i) out[9 + i] = a[i] + it[i];
}
