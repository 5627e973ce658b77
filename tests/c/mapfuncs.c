/* Functions of the coordinates for the mapping tests: each has the signature the reference's
 * cgrid_map / rgrid_map take (src/cgrid.c:793, src/rgrid.c:806).  Built into a small shared
 * library by the tests; the numpy mirrors live in tests/test_gridio.py. */
#include <complex.h>

/* x + i y and z + i (x - y): together they expose all three coordinates */
double complex map_xy(void *arg, double x, double y, double z) {
  (void)arg; (void)z;
  return x + I * y;
}

double complex map_z(void *arg, double x, double y, double z) {
  (void)arg;
  return z + I * (x - y);
}

/* a smooth complex function with a parameter read through the argument pointer */
double complex map_poly(void *arg, double x, double y, double z) {
  const double a = *(const double *)arg;
  return (1.0 + a * x - 0.2 * y + 0.1 * z + 0.05 * x * x + 0.02 * y * z) + I * (0.1 * x * y + 0.03 * z * z);
}

double rmap_poly(void *arg, double x, double y, double z) {
  const double a = *(const double *)arg;
  return 1.0 + a * x * x - 0.2 * y + 0.1 * z * y + 0.3 * z;
}

double rmap_x(void *arg, double x, double y, double z) {
  (void)arg; (void)y; (void)z;
  return x;
}

double rmap_yz(void *arg, double x, double y, double z) {
  (void)arg; (void)x;
  return 1000.0 * y + z;
}
