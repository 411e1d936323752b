/*
 * Faddeeva known-answer dump: test infrastructure only.
 *
 * Links the reference's complex_errorfunc_wofz_algo680 (src/physic.c:75-301,
 * compiled in place by oracle/Makefile) and prints, for a deterministic set of
 * points z = x + iy covering all four quadrants and the three branches of
 * TOMS 680 (power series, Taylor with continued-fraction derivatives, Laplace
 * continued fraction), the line "x y u v flag" in %.17g.
 *
 * Usage: refwofz [n_per_decade]
 */
#include <stdio.h>
#include <stdlib.h>
#include <math.h>

int complex_errorfunc_wofz_algo680(double x, double y, double u[2]);

int main(int argc, char **argv)
{
  int per = argc > 1 ? atoi(argv[1]) : 12;
  /* radii 10^-3 .. 10^2.5 in per-decade steps, 16 angles each (all quadrants),
     the axes and a few large arguments near the reflection's overflow limit */
  for (int k = 0; k <= (int)(5.5 * per); k++) {
    double r = pow(10.0, -3.0 + (double)k / per);
    for (int a = 0; a < 16; a++) {
      double th = (2.0 * M_PI * a + 0.37) / 16.0;
      double x = r * cos(th), y = r * sin(th), u[2];
      int f = complex_errorfunc_wofz_algo680(x, y, u);
      printf("%.17g %.17g %.17g %.17g %d\n", x, y, u[0], u[1], f);
    }
  }
  const double extra[][2] = {{0, 0}, {1, 0}, {-1, 0}, {0, 1}, {0, -1}, {5.3, 0}, {0, 5.3}, {1, -30}, {30, -30}, {-26.5, -1}};
  for (unsigned i = 0; i < sizeof extra / sizeof extra[0]; i++) {
    double u[2];
    int f = complex_errorfunc_wofz_algo680(extra[i][0], extra[i][1], u);
    printf("%.17g %.17g %.17g %.17g %d\n", extra[i][0], extra[i][1], u[0], u[1], f);
  }
  return 0;
}
