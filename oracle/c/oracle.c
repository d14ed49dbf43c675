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
