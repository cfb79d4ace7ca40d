// CPU accuracy check of parflow_amd/csrc/pf_math.h against long double
// references: prints the maximum error in units of the last place of the
// double result for exp, log and log1p over the argument ranges the van
// Genuchten closures use.
#include <cmath>
#include <cstdio>
#include <cstdint>
#include "pf_math.h"

static double ulp_err(double got, long double ref)
{
  double r = (double)ref;
  double u = nextafter(fabs(r), INFINITY) - fabs(r);
  return (double)(fabsl((long double)got - ref) / u);
}

int main()
{
  uint64_t s = 88172645463325252ull;
  auto rnd = [&]() {
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    return (double)(s >> 11) * (1.0 / 9007199254740992.0);
  };
  const int N = 2000000;
  double me = 0, ml = 0, m1 = 0;
  for (int q = 0; q < N; q++) {
    double x = -50.0 + 100.0 * rnd();
    me = fmax(me, ulp_err(pf_exp(x), expl((long double)x)));
    double y = pow(10.0, -12.0 + 24.0 * rnd());
    ml = fmax(ml, ulp_err(pf_log(y), logl((long double)y)));
    double t = pow(10.0, -20.0 + 32.0 * rnd());
    m1 = fmax(m1, ulp_err(pf_log1p(t), log1pl((long double)t)));
  }
  // special points
  ml = fmax(ml, ulp_err(pf_log(1.0), 0.0L) + (pf_log(1.0) != 0.0 ? 1e9 : 0.0));
  m1 = fmax(m1, (pf_log1p(0.0) != 0.0) ? 1e9 : 0.0);
  printf("exp %.3f\nlog %.3f\nlog1p %.3f\n", me, ml, m1);
  return 0;
}
