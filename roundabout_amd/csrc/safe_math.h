// safe_math.h -- position-independent exp / log / pow / erfc on the device,
// the reference's rdb_sm_* library (src/shared_module_utilities/
// rdb_safe_math_core.inc, _pow.inc, _erfc.inc): Cody-Waite reduction by
// ln 2 with a degree-13 Taylor polynomial for exp, the atanh series in
// f = (m-1)/(m+1) for log (double-double for pow), exact Dekker/Knuth
// products and sums, and rational fits for erfc.  Built without FMA
// contraction, every result is bitwise the reference's CPU value; the
// polynomial tables and the ln 2 split are the reference's
// (rdb_safe_math_consts.inc:29-128), since the bits depend on them.
#pragma once

#include <hip/hip_runtime.h>

namespace colphys {
namespace sm {

constexpr double TINY = 2.2250738585072014e-308;
constexpr double LN2_HI = 0.69314718060195446014;
constexpr double LN2_LO = -4.2009150726810845979e-11;
constexpr double INV_LN2 = 1.442695040888963387;
constexpr double EXP_MAX = 709.78271289338399679;
constexpr double EXP_MIN = -745.1332191019412;
constexpr double SQRT_HALF = 0.70710678118654757274;
constexpr double TWO54 = 18014398509481984.0;
constexpr double SPLIT = 134217729.0;
constexpr double INV_SQRTPI = 0.56418958354775627928;

__device__ __forceinline__ double inf() { return __longlong_as_double(0x7FF0000000000000LL); }
__device__ __forceinline__ double nan() { return __longlong_as_double(0x7FF8000000000000LL); }

__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

__device__ __forceinline__ void two_prod(double a, double b, double& p, double& e) {
  double c = SPLIT * a;
  double ah = c - (c - a);
  double al = a - ah;
  c = SPLIT * b;
  double bh = c - (c - b);
  double bl = b - bh;
  p = a * b;
  e = (((ah * bh - p) + ah * bl) + al * bh) + al * bl;
}

// exp(x + c) for a small correction c (rdb_sm_exp_core).
__device__ __forceinline__ double exp_core(double x, double c) {
  const double kr = round(x * INV_LN2);  // round half away from zero, as anint
  const int k = (int)kr;
  double r = x - kr * LN2_HI;
  r = r - kr * LN2_LO;
  r = r + c;
  double p = 1.0 / 6227020800.0;
  p = p * r + 1.0 / 479001600.0;
  p = p * r + 1.0 / 39916800.0;
  p = p * r + 1.0 / 3628800.0;
  p = p * r + 1.0 / 362880.0;
  p = p * r + 1.0 / 40320.0;
  p = p * r + 1.0 / 5040.0;
  p = p * r + 1.0 / 720.0;
  p = p * r + 1.0 / 120.0;
  p = p * r + 1.0 / 24.0;
  p = p * r + 1.0 / 6.0;
  p = p * r + 1.0 / 2.0;
  p = p * r + 1.0;
  p = p * r + 1.0;
  double y = (k > 1000) ? 2.0 * ldexp(p, k - 1) : ldexp(p, k);
  if (y < TINY) y = 0.0;
  return y;
}

__device__ __forceinline__ double exp(double x) {
  if (isnan(x)) return x;
  if (x > EXP_MAX) return inf();
  if (x < EXP_MIN) return 0.0;
  return exp_core(x, 0.0);
}

__device__ __forceinline__ void reduce_log(double x, double& m, int& k) {
  double xs = x;
  int kadj = 0;
  if (x < TINY) {
    xs = x * TWO54;
    kadj = -54;
  }
  int e;
  m = frexp(xs, &e);
  k = e + kadj;
  if (m < SQRT_HALF) {
    m = 2.0 * m;
    k = k - 1;
  }
}

// sum_{i=1..9} u^(i-1) / (2i+1) by Horner from the top (SM_LOG_C(9..1)).
__device__ __forceinline__ double log_series(double u) {
  double p = 1.0 / 19.0;
  p = p * u + 1.0 / 17.0;
  p = p * u + 1.0 / 15.0;
  p = p * u + 1.0 / 13.0;
  p = p * u + 1.0 / 11.0;
  p = p * u + 1.0 / 9.0;
  p = p * u + 1.0 / 7.0;
  p = p * u + 1.0 / 5.0;
  p = p * u + 1.0 / 3.0;
  return p;
}

__device__ __forceinline__ double log(double x) {
  if (isnan(x)) return x;
  if (x < 0.0) return nan();
  if (x == 0.0) return -inf();
  if (x < TINY && x * TWO54 == 0.0) return -inf();
  if (!isfinite(x)) return x;
  double m;
  int k;
  reduce_log(x, m, k);
  const double f = (m - 1.0) / (m + 1.0);
  const double u = f * f;
  const double p = log_series(u);
  const double kr = (double)k;
  const double s = kr * LN2_HI + 2.0 * f;
  const double w = ((2.0 * f) * u) * p + kr * LN2_LO;
  return s + w;
}

__device__ __forceinline__ void log_dd(double x, double& lh, double& ll) {
  if (x < TINY && x * TWO54 == 0.0) {
    lh = -inf();
    ll = 0.0;
    return;
  }
  double m, bh, bl, p1, p2, s, e;
  int k;
  reduce_log(x, m, k);
  const double a = m - 1.0;
  two_sum(m, 1.0, bh, bl);
  const double q1 = a / bh;
  two_prod(q1, bh, p1, p2);
  const double rr = ((a - p1) - p2) - q1 * bl;
  const double q2 = rr / bh;
  const double fh = q1 + q2;
  const double fl = q2 - (fh - q1);
  const double u = fh * fh;
  const double p = log_series(u);
  const double tail = ((2.0 * fh) * u) * p;
  const double kr = (double)k;
  two_sum(kr * LN2_HI, 2.0 * fh, s, e);
  lh = s;
  ll = e + ((2.0 * fl + tail) + kr * LN2_LO);
}

inline __device__ __noinline__ double pow(double x, double y) {
  if (isnan(x) || isnan(y)) return nan();
  if (y == 0.0 || x == 1.0) return 1.0;
  const double ax = fabs(x), ay = fabs(y);
  bool yint, yodd;
  if (isfinite(y)) {
    if (ay >= 9007199254740992.0) {
      yint = true;
      yodd = false;
    } else {
      yint = (ay == round(ay));
      const double h = 0.5 * ay;
      yodd = yint && (h != round(h));
    }
  } else {
    yint = false;
    yodd = false;
  }
  const bool xneg = (x < 0.0) || (x == 0.0 && signbit(x));
  double r;
  if (!isfinite(y)) {
    if (ax == 1.0) return 1.0;
    if ((ax < 1.0) == (y > 0.0)) return 0.0;
    return inf();
  }
  if (ax == 0.0) {
    r = (y > 0.0) ? 0.0 : inf();
    return (xneg && yodd) ? -r : r;
  }
  if (!isfinite(x)) {
    r = (y > 0.0) ? inf() : 0.0;
    return (xneg && yodd) ? -r : r;
  }
  if (xneg && !yint) return nan();
  double lh, ll;
  log_dd(ax, lh, ll);
  const double t = y * lh;
  if (t > 800.0) {
    r = inf();
  } else if (t < -800.0) {
    r = 0.0;
  } else {
    double ph, pl, e, c2;
    two_prod(y, lh, ph, pl);
    const double c = pl + y * ll;
    two_sum(ph, c, e, c2);
    if (e > EXP_MAX)
      r = inf();
    else if (e < EXP_MIN)
      r = 0.0;
    else
      r = exp_core(e, c2);
  }
  return (xneg && yodd) ? -r : r;
}

inline __device__ __noinline__ double erfc(double x) {
  if (isnan(x)) return x;
  const double ax = fabs(x);
  double p, q, ex;
  if (ax < 0.5) {
    const double v = x * x;
    p = -1.095865453884735107e-9;
    p = p * v + 1.472152117638980096e-8;
    p = p * v + -1.636264690374603956e-7;
    p = p * v + 1.64620383639705829e-6;
    p = p * v + -0.00001492564916290355896;
    p = p * v + 0.0001205533296923066544;
    p = p * v + -0.0008548327023363942604;
    p = p * v + 0.005223977625441806293;
    p = p * v + -0.02686617064513124181;
    p = p * v + 0.1128379167095512614;
    p = p * v + -0.376126389031837538;
    p = p * v + 1.128379167095512559;
    return 1.0 - x * p;
  }
  if (ax > 27.3) {
    ex = 0.0;
  } else {
    const double c = SPLIT * ax;
    const double xh = c - (c - ax);
    const double xl = ax - xh;
    ex = exp_core(-(xh * xh), -(xl * (ax + xh)));
  }
  if (ax < 2.0) {
    p = 4.820036176802313081e-8;
    q = 0.009303534974169982985;
    p = p * ax + 0.005247572786535052193;
    q = q * ax + 0.1013327241858480304;
    p = p * ax + 0.05719027224971907214;
    q = q * ax + 0.5046949490889137691;
    p = p * ax + 0.2819428661917610102;
    q = q * ax + 1.443014856236056964;
    p = p * ax + 0.7867094724888269441;
    q = q * ax + 2.468873583018134976;
    p = p * ax + 1.248626834329117141;
    q = q * ax + 2.377005991883625935;
    p = p * ax + 0.9999999993850857694;
    q = q * ax + 1.0;
    ex = ex * (p / q);
  } else if (ax < 4.0) {
    p = 1.000103316758091609e-9;
    q = 0.02463786793791475013;
    p = p * ax + 0.01390038004034872994;
    q = q * ax + 0.2013396690397278865;
    p = p * ax + 0.1135948493973390344;
    q = q * ax + 0.8274770365185684096;
    p = p * ax + 0.4598879010194165118;
    q = q * ax + 2.025960043609839722;
    p = p * ax + 1.086387126830294259;
    q = q * ax + 3.045871297831376889;
    p = p * ax + 1.494269944137282868;
    q = q * ax + 2.622591678871410537;
    p = p * ax + 0.9999921569124016196;
    q = q * ax + 1.0;
    ex = ex * (p / q);
  } else {
    const double rax = 1.0 / ax;
    const double w = rax * rax;
    p = -12.38208675658035851;
    q = 428.3053107970796418;
    p = p * w + -231.8021338423288;
    q = q * w + 1100.222105092567517;
    p = p * w + -273.5933562329464053;
    q = q * w + 802.2984884770494318;
    p = p * w + -95.49719132910308872;
    q = q * w + 226.1241802321382579;
    p = p * w + -12.20993252402114315;
    q = q * w + 25.91986504803799107;
    p = p * w + -0.4999999999999982792;
    q = q * w + 1.0;
    ex = ((ex * INV_SQRTPI) * rax) * (1.0 + w * (p / q));
  }
  if (ex < TINY) ex = 0.0;
  return (x < 0.0) ? 2.0 - ex : ex;
}

}  // namespace sm
}  // namespace colphys
