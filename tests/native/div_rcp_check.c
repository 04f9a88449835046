/* Checks, on the host, the division used by the pipelined Gauss-Seidel walk
 * (div_rcp2 in cfd-lite_amd/csrc/cfdl_device.hip): with y = 1/a correctly
 * rounded, q0 = x*y, r = fma(-a, q0, x) and q = fma(r, y, q0) is the
 * correctly rounded x/a (Markstein's theorem) while the quotient estimate
 * stays normal; the walk divides that by sor > 0 the same way and fixes the
 * sign of a zero result once, from q0.  It divides only with divisors in
 * [2^-64, 2^64] and dividends of frexp exponent in [-750, 750] or zero;
 * outside that it falls back to the hardware division.  Operands are drawn
 * over those whole ranges (a of either sign, sor > 0), with significands
 * near binade edges, sparse bit patterns and signed zeros; the single
 * division (with its own copysign) and the chained one must equal the
 * plain divisions bit for bit.
 * Usage: div_rcp_check SEED SAMPLES; exit status 0 when all agree. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t st;
static uint64_t next(void) {
  st ^= st << 13;
  st ^= st >> 7;
  st ^= st << 17;
  return st;
}
static double make(uint64_t mant, int ex, int neg) {
  uint64_t bits = ((uint64_t)(ex + 1023) << 52) | (mant & ((1ull << 52) - 1));
  double d;
  memcpy(&d, &bits, 8);
  return neg ? -d : d;
}
static double div_rcp(double x, double a, double y) {
  const double q0 = x * y;
  const double r = fma(-a, q0, x);
  return copysign(fma(r, y, q0), q0);
}
static double div_rcp2(double x, double a, double y, double s, double ys) {
  const double q0 = x * y;
  const double q1 = fma(fma(-a, q0, x), y, q0);
  const double p0 = q1 * ys;
  return copysign(fma(fma(-s, p0, q1), ys, p0), q0);
}
static long bad;
static void check(double x, double a, double s) {
  volatile double av = a, sv = s;
  const double ya = 1.0 / av, ys = 1.0 / sv;
  const double q = div_rcp(x, a, ya), ref = x / av;
  const double q2 = div_rcp2(x, a, ya, s, ys), ref2 = ref / sv;
  if (memcmp(&q, &ref, 8) != 0 || memcmp(&q2, &ref2, 8) != 0) {
    if (bad < 5) printf("mismatch x=%a a=%a s=%a: %a %a vs %a %a\n", x, a, s, q, q2, ref, ref2);
    bad++;
  }
}
static uint64_t mant(long i) {
  switch (i & 7) {
    case 1: return next() & 0xffff;    /* just above a power of two */
    case 2: return ~(next() & 0xffff); /* just below a power of two */
    case 3: return next() & 0xff;
    case 4: return ~(next() & 0xff);
    case 5: return next() >> (next() & 63);
    case 6: return next() << (next() & 31);
    case 7: return 0;
    default: return next();
  }
}
int main(int argc, char **argv) {
  if (argc < 3) return 2;
  st = 0x9E3779B97F4A7C15ull ^ (uint64_t)atol(argv[1]) * 0x2545F4914F6CDD1Dull;
  const long n = atol(argv[2]);
  for (long i = 0; i < n; i++) {
    /* frexp exponent e means |x| in [2^(e-1), 2^e): unbiased exponent e-1 */
    const int ex = (int)(next() % 1501) - 751, ea = (int)(next() % 129) - 64;
    const int es = (int)(next() % 129) - 64;
    const double x = make(mant(i), ex, (int)(next() & 1));
    const double a = make(mant(i >> 3), ea == 64 ? 63 : ea, (int)(next() & 1));
    const double s = make(mant(i >> 6), es == 64 ? 63 : es, 0);
    check(x, a, s);
    check(x, a, 1.02);
    check(make(next(), (int)(next() % 120) - 60, (int)(next() & 1)), a, 1.0);
  }
  const double edges[] = {0x1p-64, -0x1p-64, 0x1p64, -0x1p64, 1.02, 3.0, 0x1.fffffffffffffp63};
  for (int k = 0; k < 7; k++) {
    check(0.0, edges[k], 1.02);
    check(-0.0, edges[k], 1.02);
    check(0x1.fffffffffffffp749, edges[k], 0x1p-64);
    check(-0x1p-751, edges[k], 0x1p64);
    check(0x1.8p-751, edges[k], 0x1.fffffffffffffp63);
  }
  printf("samples %ld mismatches %ld\n", 3 * n + 35, bad);
  return bad != 0;
}
