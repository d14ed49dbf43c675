/*
 * FP64 all-pairs accelerations with Neumaier-compensated summation -- TEST INFRASTRUCTURE ONLY
 * (the oracle; linked only by tests and the benchmark's CPU-baseline leg).
 *
 * Same formula as the NumPy oracle (oracle/oracle.py: accel_f64), i.e. the reference's pair
 * formula (src/gravitation/kernel/py1.py:57-64) in the all-pairs form with G factored out
 * (src/gravitation/kernel/pc2.py:60-90), plus optional Plummer softening eps:
 *   a_i = G * sum_{j != i} m_j (r_j - r_i) (|r_j - r_i|^2 + eps^2)^(-3/2)
 * pos is (n, 3) row-major, m is (n), out is (nrows, 3) for the listed rows.
 */
#include <math.h>
#include <stddef.h>

/* Neumaier's variant of Kahan summation: s + c carries the running sum. */
static inline void neumaier_add(double *s, double *c, double x) {
  double t = *s + x;
  if (fabs(*s) >= fabs(x))
    *c += (*s - t) + x;
  else
    *c += (x - t) + *s;
  *s = t;
}

void oracle_accel_f64(const double *pos, const double *m, long n, double G, double eps,
                      const long *rows, long nrows, double *out) {
  const double eps2 = eps * eps;
  long r;
#pragma omp parallel for schedule(dynamic, 16)
  for (r = 0; r < nrows; ++r) {
    const long i = rows ? rows[r] : r;
    const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
    double sx = 0, sy = 0, sz = 0, cx = 0, cy = 0, cz = 0;
    for (long j = 0; j < n; ++j) {
      if (j == i && eps2 == 0.0) continue;
      const double dx = pos[3 * j] - xi, dy = pos[3 * j + 1] - yi, dz = pos[3 * j + 2] - zi;
      const double r2 = dx * dx + dy * dy + dz * dz + eps2;
      const double w = m[j] / (r2 * sqrt(r2));
      neumaier_add(&sx, &cx, dx * w);
      neumaier_add(&sy, &cy, dy * w);
      neumaier_add(&sz, &cz, dz * w);
    }
    out[3 * r] = G * (sx + cx);
    out[3 * r + 1] = G * (sy + cy);
    out[3 * r + 2] = G * (sz + cz);
  }
}

/*
 * Vector-Jacobian products of the accelerations above with a cotangent mu (n, 3), in FP64 --
 * the same formula as the NumPy oracle's vjp_accel_f64 (oracle/oracle.py):
 *   (J_r^T mu)_k = G sum_{j != k} T_kj (m_k mu_j - m_j mu_k),  T_kj = s (I - 3 d d^T / q)
 *   (J_m^T mu)_k = -G sum_{j != k} s_kj d_kj . mu_j
 * with d_kj = r_j - r_k, q = |d|^2 + eps^2, s = q^(-3/2).  jr is (n, 3), jm is (n).
 */
void oracle_vjp_f64(const double *pos, const double *m, const double *mu, long n, double G,
                    double eps, double *jr, double *jm) {
  const double eps2 = eps * eps;
  long k;
#pragma omp parallel for schedule(dynamic, 16)
  for (k = 0; k < n; ++k) {
    const double xk = pos[3 * k], yk = pos[3 * k + 1], zk = pos[3 * k + 2], mk = m[k];
    const double ux = mu[3 * k], uy = mu[3 * k + 1], uz = mu[3 * k + 2];
    double sx = 0, sy = 0, sz = 0, cx = 0, cy = 0, cz = 0, sm = 0, cm = 0;
    for (long j = 0; j < n; ++j) {
      if (j == k && eps2 == 0.0) continue;
      const double dx = pos[3 * j] - xk, dy = pos[3 * j + 1] - yk, dz = pos[3 * j + 2] - zk;
      const double q = dx * dx + dy * dy + dz * dz + eps2;
      const double s = 1.0 / (q * sqrt(q));
      const double wx = mk * mu[3 * j] - m[j] * ux;
      const double wy = mk * mu[3 * j + 1] - m[j] * uy;
      const double wz = mk * mu[3 * j + 2] - m[j] * uz;
      const double t = 3.0 * (dx * wx + dy * wy + dz * wz) / q;
      neumaier_add(&sx, &cx, s * (wx - t * dx));
      neumaier_add(&sy, &cy, s * (wy - t * dy));
      neumaier_add(&sz, &cz, s * (wz - t * dz));
      neumaier_add(&sm, &cm, s * (dx * mu[3 * j] + dy * mu[3 * j + 1] + dz * mu[3 * j + 2]));
    }
    jr[3 * k] = G * (sx + cx);
    jr[3 * k + 1] = G * (sy + cy);
    jr[3 * k + 2] = G * (sz + cz);
    jm[k] = -G * (sm + cm);
  }
}
