// Host-side setup of the timestep: the Gaussian lookup table, the reaction
// rate -> binding/unbinding radius solver, the reaction superstructure and
// per-reaction parameters, and the dcmt stream parameters and seeds.
//
// These run once per simulation on the CPU, in the same precision as the
// reference (FP32 for the table and smolreact, FP64 for rxnparam), and are
// checked bit for bit against the reference's own compiled code.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "smd_setup.h"

namespace smd {

// ---------------------------------------------------------------------------
// Gaussian table: stats.cpp:69-282.  The quantile Phi^-1((i + 1/2)/n) is found
// by a 30-step FP32 bisection on erf computed from the regularised incomplete
// gamma function P(1/2, x^2) (series below a+1, Lentz-style continued fraction
// above; Numerical Recipes' method and Lanczos log-gamma coefficients).
// ---------------------------------------------------------------------------
namespace {

float log_gamma_f(float x) {
  // stats.cpp:69-107 (only the branches reachable from P(1/2, .) matter,
  // but the function is complete).
  static const float lanczos[6] = {76.18009173f, -86.50532033f, 24.01409822f,
                                   -1.231739516f, 0.120858003e-2f, -0.536382e-5f};
  const float half_log_pi = 0.572364942f;  // ln(sqrt(pi)) = ln Gamma(1/2)
  float acc, t;
  if (x == std::floor(x) && x <= 0) return std::numeric_limits<float>::max();
  if (x == std::floor(x)) {
    acc = 0;
    for (t = 2; t < x - 0.1f; t += 1) acc += std::log(t);
    return acc;
  }
  if (x == 0.5f) return half_log_pi;
  if (2 * x == std::floor(2 * x) && x > 0) {
    acc = half_log_pi;
    for (t = 0.5; (double)t < (double)x - 0.1; t += 1) acc += std::log(t);
    return acc;
  }
  if (2 * x == std::floor(2 * x)) {
    acc = half_log_pi;
    for (t = 0.5; (double)t < -(double)x + 0.1; t += 1) acc -= std::log(t);
    return acc;
  }
  if (x < 0) return log_gamma_f(x + 1) - std::log(-x);
  x -= 1.0f;
  t = x + 5.5f;
  t -= (x + 0.5f) * std::log(t);
  acc = 1.0f;
  for (int j = 0; j <= 5; ++j) {
    x += 1.0f;
    acc += lanczos[j] / x;
  }
  return -t + std::log(2.50662827465f * acc);
}

float gamma_p_f(float a, float x) {
  // stats.cpp:110-154
  const float eps = 3e-7f;
  if (x < 0 || a <= 0) return -1;
  if (x == 0) return 0;
  if (x < a + 1) {
    float ap = a, term = 1 / a, sum = 1 / a;
    while (std::fabs(term) > std::fabs(sum) * eps && ap - a < 100) {
      ap += 1;
      term *= x / ap;
      sum += term;
    }
    return sum * std::exp(-x + a * std::log(x) - log_gamma_f(a));
  }
  float gold = 0, g = 1, fac = 1, b1 = 1, b0 = 0, a0 = 1, a1 = x;
  for (float an = 1; an < 100; an++) {
    const float ana = an - a;
    a0 = (a1 + a0 * ana) * fac;
    b0 = (b1 + b0 * ana) * fac;
    const float anf = an * fac;
    a1 = x * a0 + anf * a1;
    b1 = x * b0 + anf * b1;
    if (a1) {
      fac = 1.0f / a1;
      g = b1 * fac;
      if (std::fabs((g - gold) / g) < eps) return 1.0f - std::exp(-x + a * std::log(x) - log_gamma_f(a)) * g;
      gold = g;
    }
  }
  return -1;
}

float erf_f(float x) { return x < 0 ? -gamma_p_f(0.5f, x * x) : gamma_p_f(0.5f, x * x); }

// stats.cpp:161-180: bisection of a monotone function on [x1, x2].
float bisect_inverse(float (*fn)(float), float y, float x1, float x2, int n) {
  float dx;
  if (fn(x1) < fn(x2)) {
    dx = x2 - x1;
  } else {
    dx = x1 - x2;
    x1 = x2;
  }
  for (; n > 0; n--) {
    dx *= 0.5f;
    x2 = x1 + dx;
    if (fn(x2) < y) x1 = x2;
  }
  return x1 + 0.5f * dx;
}

}  // namespace

void gauss_table(float* a, int n) {
  // randtableF(a, n, 1) (stats.cpp:202-211)
  const float sqrt2 = 1.41421356237f;
  const float dy = 2.0f / n;
  for (int i = 0; i < n / 2; ++i) a[i] = sqrt2 * bisect_inverse(erf_f, (i + 0.5f) * dy - 1.0f, -20, 20, 30);
  for (int i = n / 2; i < n; ++i) a[i] = -a[n - i - 1];
  // randshuffletableF (stats.cpp:220-232) swaps a[i] with a[intrand(n)].  The
  // generator behind intrand is a static mt_struct that is never configured
  // (stats.cpp:44-57: nn = 0, all masks 0), so every draw is 0 and the
  // "shuffle" is the sequence of swaps a[i] <-> a[0]: a right rotation by one.
  for (int i = 0; i < n; ++i) {
    const float t = a[i];
    a[i] = a[0];
    a[0] = t;
  }
}

// ---------------------------------------------------------------------------
// rxnparam (rxnparam.cpp:42-322; Andrews & Bray, Phys. Biol. 2004).  The
// tabulated reduced rates are data of that method: yi for irreversible
// reactions (31 reduced steps), yr for reversible with b > a (16 x 31),
// yb for b < a (11 x 31, -1 = not computed).  Lookups use 4-point Lagrange
// interpolation in ln(s/a) and in b/a (or ln(b/a)).
// ---------------------------------------------------------------------------
namespace {

const double kPi = 3.14159265358979323846;

const double kYi[31] = {
    4.188584, 4.188381, 4.187971, 4.187141, 4.185479, 4.182243, 4.1761835, 4.1652925,
    4.146168, 4.112737, 4.054173, 3.95406,  3.7899875, 3.536844, 3.178271,  2.7239775,
    2.2178055, 1.7220855, 1.2875175, 0.936562, 0.668167, 0.470021, 0.327102, 0.225858,
    0.1551475, 0.1061375, 0.072355, 0.049183, 0.0333535, 0.022576, 0.015257};

const double kYr[16 * 31] = {
    4.188790, 4.188790, 4.188790, 4.188788, 4.188785, 4.188775, 4.188747, 4.188665, 4.188438, 4.187836, 4.186351,
    4.182961, 4.175522, 4.158666, 4.119598, 4.036667, 3.883389, 3.641076, 3.316941, 2.945750, 2.566064, 2.203901,
    1.872874, 1.578674, 1.322500, 1.103003, 0.916821, 0.759962, 0.628537, 0.518952, 0.428136,
    4.188790, 4.188790, 4.188789, 4.188786, 4.188778, 4.188756, 4.188692, 4.188509, 4.188004, 4.186693, 4.183529,
    4.176437, 4.160926, 4.125854, 4.047256, 3.889954, 3.621782, 3.237397, 2.773578, 2.289205, 1.831668, 1.427492,
    1.087417, 0.811891, 0.595737, 0.430950, 0.308026, 0.217994, 0.152981, 0.106569, 0.073768,
    4.188790, 4.188789, 4.188788, 4.188783, 4.188768, 4.188727, 4.188608, 4.188272, 4.187364, 4.185053, 4.179623,
    4.167651, 4.141470, 4.082687, 3.956817, 3.722364, 3.355690, 2.877999, 2.353690, 1.850847, 1.411394, 1.050895,
    0.768036, 0.553077, 0.393496, 0.277134, 0.193535, 0.134203, 0.092515, 0.063465, 0.043358,
    4.188790, 4.188789, 4.188786, 4.188777, 4.188753, 4.188682, 4.188480, 4.187919, 4.186431, 4.182757, 4.174363,
    4.156118, 4.116203, 4.028391, 3.851515, 3.546916, 3.110183, 2.587884, 2.055578, 1.574898, 1.174608, 0.858327,
    0.617223, 0.438159, 0.307824, 0.214440, 0.148358, 0.102061, 0.069885, 0.047671, 0.032414,
    4.188789, 4.188788, 4.188783, 4.188769, 4.188729, 4.188614, 4.188288, 4.187399, 4.185108, 4.179638, 4.167499,
    4.141379, 4.084571, 3.964576, 3.739474, 3.381771, 2.906433, 2.372992, 1.854444, 1.401333, 1.032632, 0.746509,
    0.531766, 0.374442, 0.261247, 0.180929, 0.124557, 0.085332, 0.058229, 0.039604, 0.026865,
    4.188789, 4.188786, 4.188778, 4.188756, 4.188692, 4.188510, 4.188002, 4.186650, 4.183288, 4.175566, 4.158864,
    4.123229, 4.047257, 3.896024, 3.632545, 3.242239, 2.750962, 2.220229, 1.717239, 1.285681, 0.939696, 0.674540,
    0.477621, 0.334612, 0.232460, 0.160415, 0.110103, 0.075242, 0.051236, 0.034789, 0.023565,
    4.188788, 4.188784, 4.188772, 4.188737, 4.188637, 4.188354, 4.187585, 4.185607, 4.180895, 4.170490, 4.148460,
    4.102035, 4.006907, 3.829897, 3.541151, 3.133335, 2.635739, 2.109597, 1.619284, 1.204298, 0.875185, 0.625190,
    0.440882, 0.307818, 0.213235, 0.146800, 0.100560, 0.068608, 0.046655, 0.031645, 0.021415,
    4.188787, 4.188780, 4.188761, 4.188707, 4.188552, 4.188124, 4.186995, 4.184222, 4.177931, 4.164527, 4.136619,
    4.079198, 3.967945, 3.772932, 3.468711, 3.050092, 2.548726, 2.026932, 1.547039, 1.144954, 0.828592, 0.589866,
    0.414774, 0.288895, 0.199728, 0.137273, 0.093903, 0.063994, 0.043478, 0.029467, 0.019931,
    4.188785, 4.188774, 4.188745, 4.188661, 4.188426, 4.187795, 4.186203, 4.182494, 4.174471, 4.157859, 4.123939,
    4.056910, 3.934028, 3.727335, 3.412145, 2.985387, 2.481620, 1.963935, 1.492519, 1.100533, 0.793976, 0.563797,
    0.395615, 0.275071, 0.189896, 0.130359, 0.089086, 0.060661, 0.041187, 0.027899, 0.018863,
    4.188782, 4.188766, 4.188720, 4.188592, 4.188244, 4.187347, 4.185205, 4.180486, 4.170691, 4.150875, 4.111574,
    4.037574, 3.906673, 3.691217, 3.367270, 2.934386, 2.429282, 1.915212, 1.450660, 1.066615, 0.767733, 0.544143,
    0.381225, 0.264724, 0.182558, 0.125209, 0.085504, 0.058187, 0.039490, 0.026741, 0.018074,
    4.188777, 4.188752, 4.188683, 4.188492, 4.187993, 4.186777, 4.184049, 4.178342, 4.166861, 4.144167, 4.100756,
    4.021817, 3.884852, 3.662169, 3.331386, 2.893934, 2.388069, 1.877097, 1.418051, 1.040351, 0.747549, 0.529083,
    0.370239, 0.256844, 0.176983, 0.121304, 0.082791, 0.056318, 0.038211, 0.025871, 0.017486,
    4.188770, 4.188732, 4.188628, 4.188352, 4.187671, 4.186115, 4.182830, 4.176234, 4.163261, 4.138284, 4.092055,
    4.009206, 3.867171, 3.638731, 3.302580, 2.861651, 2.355385, 1.846979, 1.392364, 1.019841, 0.731849, 0.517409,
    0.361743, 0.250766, 0.172684, 0.118295, 0.080706, 0.054883, 0.037230, 0.025204, 0.017035,
    4.188759, 4.188702, 4.188551, 4.188171, 4.187294, 4.185421, 4.181657, 4.174292, 4.160089, 4.133434, 4.084980,
    3.998932, 3.852801, 3.619743, 3.279339, 2.835771, 2.329266, 1.822929, 1.372041, 1.003734, 0.719559, 0.508296,
    0.355130, 0.246037, 0.169346, 0.115966, 0.079096, 0.053783, 0.036482, 0.024699, 0.016694,
    4.188742, 4.188659, 4.188449, 4.187958, 4.186900, 4.184765, 4.180606, 4.172596, 4.157440, 4.129551, 4.079183,
    3.990499, 3.841031, 3.604270, 3.260519, 2.814871, 2.308170, 1.803653, 1.355971, 0.991032, 0.709899, 0.501150,
    0.349948, 0.242337, 0.166741, 0.114156, 0.077855, 0.052940, 0.035912, 0.024316, 0.016437,
    4.188719, 4.188603, 4.188330, 4.187736, 4.186534, 4.184200, 4.179711, 4.171172, 4.155254, 4.126372, 4.074457,
    3.983648, 3.831433, 3.591609, 3.245083, 2.797662, 2.290945, 1.788288, 1.343237, 0.981003, 0.702295, 0.495533,
    0.345879, 0.239437, 0.164709, 0.112756, 0.076902, 0.052295, 0.035478, 0.024023, 0.016239,
    4.188688, 4.188536, 4.188204, 4.187532, 4.186227, 4.183725, 4.178948, 4.169962, 4.153461, 4.123737, 4.070508,
    3.977891, 3.823388, 3.580958, 3.232035, 2.783281, 2.277011, 1.776032, 1.333144, 0.973094, 0.696303, 0.491107,
    0.342676, 0.237173, 0.163143, 0.111689, 0.076180, 0.051808, 0.035151, 0.023802, 0.016089};

const double kYb[11 * 31] = {
    4.188790, 4.188791, 4.188791, 4.188793, 4.188798, 4.188814, 4.188859, 4.188986, 4.189328, 4.190189, 4.192213,
    4.196874, 4.208210, 4.236983, 4.307724, 4.473512, 4.852754, 5.723574, 7.838600, 13.880005, 40.362527, -1,
    -1, -1, -1, -1, -1, -1, -1, -1, -1,
    4.188790, 4.188791, 4.188791, 4.188793, 4.188798, 4.188813, 4.188858, 4.188982, 4.189319, 4.190165, 4.192156,
    4.196738, 4.207879, 4.236133, 4.305531, 4.467899, 4.838139, 5.683061, 7.710538, 13.356740, 36.482501, -1,
    -1, -1, -1, -1, -1, -1, -1, -1, -1,
    4.188790, 4.188791, 4.188791, 4.188793, 4.188798, 4.188812, 4.188854, 4.188973, 4.189292, 4.190095, 4.191984,
    4.196332, 4.206887, 4.233594, 4.298997, 4.451212, 4.795122, 5.566006, 7.352764, 11.994360, 28.085852, 261.834153,
    -1, -1, -1, -1, -1, -1, -1, -1, -1,
    4.188790, 4.188791, 4.188791, 4.188792, 4.188797, 4.188810, 4.188848, 4.188956, 4.189246, 4.189978, 4.191699,
    4.195657, 4.205243, 4.229397, 4.288214, 4.424093, 4.726265, 5.384525, 6.831926, 10.241985, 19.932261, 69.581522,
    -1, -1, -1, -1, -1, -1, -1, -1, -1,
    4.188790, 4.188791, 4.188791, 4.188792, 4.188796, 4.188807, 4.188840, 4.188933, 4.189183, 4.189814, 4.191300,
    4.194716, 4.202957, 4.223593, 4.273478, 4.387310, 4.635278, 5.155725, 6.228070, 8.494834, 13.840039, 30.707995,
    217.685066, -1, -1, -1, -1, -1, -1, -1, -1,
    4.188790, 4.188791, 4.188791, 4.188792, 4.188795, 4.188803, 4.188829, 4.188903, 4.189101, 4.189604, 4.190789,
    4.193514, 4.200048, 4.216255, 4.254956, 4.342131, 4.526936, 4.897901, 5.610205, 6.966144, 9.704274, 16.182908,
    37.778925, 387.192994, -1, -1, -1, -1, -1, -1, -1,
    4.188790, 4.188790, 4.188791, 4.188791, 4.188793, 4.188799, 4.188817, 4.188866, 4.189002, 4.189348, 4.190168,
    4.192054, 4.196535, 4.207468, 4.233117, 4.289818, 4.406055, 4.627980, 5.025468, 5.719284, 6.973164, 9.454673,
    15.165237, 32.669841, 175.505441, -1, -1, -1, -1, -1, -1,
    4.188790, 4.188790, 4.188790, 4.188791, 4.188791, 4.188794, 4.188801, 4.188823, 4.188885, 4.189047, 4.189439,
    4.190344, 4.192444, 4.197336, 4.208277, 4.231803, 4.277545, 4.359611, 4.499274, 4.738728, 5.168746, 5.973516,
    7.554064, 10.974866, 20.015747, 59.729701, -1, -1, -1, -1, -1,
    4.188790, 4.188790, 4.188790, 4.188790, 4.188789, 4.188788, 4.188784, 4.188774, 4.188751, 4.188701, 4.188602,
    4.188390, 4.187803, 4.185972, 4.180908, 4.169592, 4.145768, 4.102670, 4.041148, 3.980771, 3.961901, 4.039620,
    4.292076, 4.858676, 6.044956, 8.672140, 15.704344, 48.437802, -1, -1, -1,
    4.188790, 4.188790, 4.188790, 4.188789, 4.188787, 4.188781, 4.188764, 4.188718, 4.188599, 4.188311, 4.187661,
    4.186201, 4.182644, 4.173501, 4.151495, 4.104610, 4.014334, 3.863341, 3.650600, 3.398722, 3.141206, 2.906174,
    2.713737, 2.580237, 2.524103, 2.574058, 2.785372, 3.278275, 4.350631, 6.897568, 14.924070,
    4.188790, 4.188790, 4.188790, 4.188788, 4.188784, 4.188773, 4.188742, 4.188656, 4.188430, 4.187878, 4.186617,
    4.183786, 4.177000, 4.160053, 4.120473, 4.038061, 3.886243, 3.645077, 3.322103, 2.951900, 2.573084, 2.212134,
    1.883264, 1.592189, 1.339766, 1.124235, 0.942582, 0.791354, 0.667141, 0.566827, 0.487862};

// Grid of the tables: ln(s/a) from 3 down in steps of 0.2 (31 points); b/a
// from 0 in steps of 0.1 (11 points); ln(b/a) from 0 in steps of 0.2 (16).
const double kSLo = 3, kSInc = -0.2;
const double kBLoR = 0, kBIncR = 0.2;
const double kBLoB = 0, kBIncB = 0.1;
const int kSNum = 31, kBNumR = 16, kBNumB = 11;

// Lagrange weights of the 4 nodes x[0..3] (spacing h) at position v.
void lagrange4(double v, const double x[4], double h, double w[4]) {
  w[0] = (v - x[1]) * (v - x[2]) * (v - x[3]) / (-6.0 * h * h * h);
  w[1] = (v - x[0]) * (v - x[2]) * (v - x[3]) / (2.0 * h * h * h);
  w[2] = (v - x[0]) * (v - x[1]) * (v - x[3]) / (-2.0 * h * h * h);
  w[3] = (v - x[0]) * (v - x[1]) * (v - x[2]) / (6.0 * h * h * h);
}

void step_nodes(int sidx, double x[4]) {
  for (int i = 0; i < 4; i++) x[i] = kSLo + (sidx - 1 + i) * kSInc;
}


}  // namespace

double erfc_d(double x) {
  // Numerical Recipes erfcc (rxnparam.cpp:316-322)
  const double z = std::fabs(x);
  const double t = 1.0 / (1.0 + 0.5 * z);
  const double ans =
      t * std::exp(-z * z - 1.26551223 +
                   t * (1.00002368 +
                        t * (0.37409196 +
                             t * (0.09678418 +
                                  t * (-0.18628806 +
                                       t * (0.27886807 +
                                            t * (-1.13520398 + t * (1.48851587 + t * (-0.82215223 + t * 0.17087277)))))))));
  return x >= 0.0 ? ans : 2.0 - ans;
}

double numrxnrate(double step, double a, double b) {
  if (step < 0 || a < 0) return -1;
  if (step == 0 && b >= 0 && b < 1) return -1;
  if (step == 0) return 0;
  const double ls = std::log(step / a);
  b /= a;
  int sidx = (int)((ls - kSLo) / kSInc);
  double x[4], w[4], y[4];
  step_nodes(sidx, x);
  lagrange4(ls, x, kSInc, w);
  double ans;
  if (b < 0) {
    ans = 0;
    for (int i = 0; i < 4; i++) {
      const int k = sidx - 1 + i;
      if (k >= kSNum) ans += w[i] * 2.0 * kPi * std::exp(2.0 * x[i]);
      else if (k < 0) ans += w[i] * 4.0 * kPi / 3.0;
      else ans += w[i] * kYi[k];
    }
  } else if (b < 1.0) {
    int bidx = (int)((b - kBLoB) / kBIncB);
    if (bidx < 1) bidx = 1;
    else if (bidx + 2 >= kBNumB) bidx = kBNumB - 3;
    while (sidx + 3 >= kSNum || (sidx > 0 && kYb[(bidx - 1) * kSNum + sidx + 3] < 0)) sidx--;
    step_nodes(sidx, x);
    lagrange4(ls, x, kSInc, w);
    for (int j = 0; j < 4; j++) {
      y[j] = 0;
      for (int i = 0; i < 4; i++) {
        const int k = sidx - 1 + i;
        if (k >= kSNum) y[j] += w[i] * kYb[(bidx - 1 + j) * kSNum];
        else if (k < 0) y[j] += w[i] * 4.0 * kPi / 3.0;
        else y[j] += w[i] * kYb[(bidx - 1 + j) * kSNum + k];
      }
    }
    double xb[4], wb[4];
    for (int j = 0; j < 4; j++) xb[j] = kBLoB + (bidx - 1 + j) * kBIncB;
    lagrange4(b, xb, kBIncB, wb);
    ans = wb[0] * y[0] + wb[1] * y[1] + wb[2] * y[2] + wb[3] * y[3];
  } else {
    const double lb = std::log(b);
    int bidx = (int)((lb - kBLoR) / kBIncR);
    if (bidx < 1) bidx = 1;
    else if (bidx + 2 >= kBNumR) bidx = kBNumR - 3;
    for (int j = 0; j < 4; j++) {
      y[j] = 0;
      for (int i = 0; i < 4; i++) {
        const int k = sidx - 1 + i;
        if (k >= kSNum && lb == 0) y[j] += w[i] * 2.0 * kPi * std::exp(x[i]) * (1.0 + std::exp(x[i]));
        else if (k >= kSNum) y[j] += w[i] * 2.0 * kPi * std::exp(2.0 * x[i]) * std::exp(lb) / (std::exp(lb) - 1.0);
        else if (k < 0) y[j] += w[i] * 4.0 * kPi / 3.0;
        else y[j] += w[i] * kYr[(bidx - 1 + j) * kSNum + k];
      }
    }
    double xb[4], wb[4];
    for (int j = 0; j < 4; j++) xb[j] = kBLoR + (bidx - 1 + j) * kBIncR;
    lagrange4(lb, xb, kBIncR, wb);
    ans = wb[0] * y[0] + wb[1] * y[1] + wb[2] * y[2] + wb[3] * y[3];
  }
  return ans * a * a * a;
}

double actrxnrate(double step, double a) {
  // rxnparam.cpp:231-239: rate with a flat radial distribution beyond a.
  const double sqrt2pi = 2.50662827462;
  if (step < 0 || a <= 0) return -1;
  if (step == 0) return 0;
  step /= a;
  double ka = 4.0 * kPi / 3.0 * (erfc_d(std::sqrt(2.0) / step) + step * std::sqrt(2.0 / kPi));
  ka += 2.0 * sqrt2pi / 3.0 * step * (step * step - 1.0) * (std::exp(-2.0 / step / step) - 1.0);
  return ka * a * a * a;
}

double bindingradius(double rate, double dt, double difc, double b, int rel) {
  // rxnparam.cpp:253-277: numrxnrate grows monotonically with a, so double
  // a from the rms step until the target rate*dt is passed, then bisect 15x.
  if (rate < 0 || dt < 0 || difc <= 0) return -1;
  if (rate == 0) return 0;
  if (dt == 0) {
    if (b < 0) return rate / (4 * kPi * difc);
    if (rel && b > 1) return rate * (1 - 1 / b) / (4 * kPi * difc);
    if (rel && b <= 1) return -1;
    if (b > 0) return rate / (4 * kPi * difc + rate / b);
    return -1;
  }
  const double step = std::sqrt(2.0 * difc * dt);
  double lo = 0, a = step;
  while (numrxnrate(step, a, rel ? b * a : b) < rate * dt) {
    lo = a;
    a *= 2.0;
  }
  double dif = a - lo;
  for (int n = 0; n < 15; n++) {
    dif *= 0.5;
    a = lo + dif;
    if (numrxnrate(step, a, rel ? b * a : b) < rate * dt) lo = a;
  }
  return lo + 0.5 * dif;
}

double unbindingradius(double pgem, double dt, double difc, double a) {
  // rxnparam.cpp:287-308: geminate rebinding probability 1 - k_irr/k_rev(b)
  // decreases with b; double then bisect 15x.  Returns -(max pgem) when the
  // requested pgem cannot be reached, -2 for illegal input.
  if (pgem <= 0 || pgem >= 1 || difc <= 0 || a <= 0 || dt < 0) return -2;
  if (dt == 0) return a / pgem;
  const double step = std::sqrt(2.0 * difc * dt);
  const double ki = numrxnrate(step, a, -1);
  const double kmax = numrxnrate(step, a, 0);
  if (1.0 - ki / kmax < pgem) return ki / kmax - 1.0;
  double lo = 0, b = a;
  while (1.0 - ki / numrxnrate(step, a, b) > pgem) {
    lo = b;
    b *= 2.0;
  }
  double dif = b - lo;
  for (int n = 0; n < 15; n++) {
    dif *= 0.5;
    b = lo + dif;
    if (1.0 - ki / numrxnrate(step, a, b) > pgem) lo = b;
  }
  return lo + 0.5 * dif;
}

// ---------------------------------------------------------------------------
// Reaction superstructure and per-reaction parameters (smolreact.cpp).
// ---------------------------------------------------------------------------
namespace {

struct OrderCsr {
  int perOrder = 0;
  std::vector<int> count, start, list, indexes;
};

void build_csr(const smd_sim_spec& sp, OrderCsr csr[3]) {
  // fillReactionOrderInfo (smolreact.cpp:173-271): per order, reactions
  // listed per reactant type in global index order; an order-2 reaction is
  // listed under both reactants.
  const int T = sp.num_types;
  for (int o = 0; o < 3; ++o) {
    csr[o].perOrder = 0;
    csr[o].count.assign(T, 0);
    csr[o].start.assign(T, 0);
    csr[o].indexes.clear();
  }
  for (int r = 0; r < sp.num_reactions; ++r) {
    const smd_reaction& R = sp.rxn[r];
    const int o = R.order;
    csr[o].indexes.push_back(r);
    csr[o].perOrder++;
    if (o > 0) {
      csr[o].count[R.reactants[0]]++;
      if (o == 2) csr[o].count[R.reactants[1]]++;
    }
  }
  for (int o = 0; o < 3; ++o) {
    for (int t = 1; t < T; ++t) csr[o].start[t] = csr[o].start[t - 1] + csr[o].count[t - 1];
    const int total = csr[o].start[T - 1] + csr[o].count[T - 1];
    csr[o].list.assign(total, 0);
    std::vector<int> fill = csr[o].start;
    for (int r = 0; r < sp.num_reactions; ++r) {
      const smd_reaction& R = sp.rxn[r];
      if (R.order != o || o == 0) continue;
      csr[o].list[fill[R.reactants[0]]++] = r;
      if (o == 2) csr[o].list[fill[R.reactants[1]]++] = r;
    }
  }
}

int prod_x(const smd_reaction& R) { return R.nprod >= 1 ? R.products[0] : -1; }
int prod_y(const smd_reaction& R) { return R.nprod >= 2 ? R.products[1] : -1; }

// findreverserxn (smolreact.cpp:73-146)
int find_reverse(const smd_sim_spec& sp, const OrderCsr csr[3], int rxn) {
  const smd_reaction& R = sp.rxn[rxn];
  const int order = R.order;
  if (order == 0 || R.nprod < 1 || csr[R.nprod].perOrder == 0) return -1;
  const int ro = R.nprod;
  const int p0 = R.products[0];
  const OrderCsr& c = csr[ro];
  for (int j = 0; j < c.count[p0]; ++j) {
    const int g = c.list[c.start[p0] + j];
    const smd_reaction& G = sp.rxn[g];
    if (ro > 1) {
      const bool same = G.reactants[0] == R.products[0] && G.reactants[1] == R.products[1];
      const bool swapped = G.reactants[0] == R.products[1] && G.reactants[1] == R.products[0];
      if (!same && !swapped) continue;
    }
    if (order == 1) {
      if (G.products[0] == R.reactants[0]) return g;
    } else {
      const bool same = G.products[0] == R.reactants[0] && G.products[1] == R.reactants[1];
      const bool swapped = G.products[0] == R.reactants[1] && G.products[1] == R.reactants[0];
      if (same || swapped) return g;
    }
  }
  return -1;
}

}  // namespace

bool derive_tables(const smd_sim_spec& sp, ReplicaTables& T, smd_rxn_derived* der, std::string& err) {
  std::memset(&T, 0, sizeof(T));
  const int NT = sp.num_types, NR = sp.num_reactions;
  if (NT < 2 || NT > kMaxTypes) { err = "num_types must be in [2, 16]"; return false; }
  if (NR < 0 || NR > kMaxRxn) { err = "num_reactions must be in [0, 32]"; return false; }
  if (!(sp.dt > 0)) { err = "dt must be > 0"; return false; }
  for (int r = 0; r < NR; ++r) {
    const smd_reaction& R = sp.rxn[r];
    if (R.order < 0 || R.order > 2 || R.nprod < 0 || R.nprod > 2) { err = "bad order/nprod"; return false; }
    for (int k = 0; k < R.order; ++k)
      if (R.reactants[k] < 1 || R.reactants[k] >= NT) { err = "reactant id out of range"; return false; }
    for (int k = 0; k < R.nprod; ++k)
      if (R.products[k] < 1 || R.products[k] >= NT) { err = "product id out of range"; return false; }
  }
  T.numTypes = NT;
  T.numRxn = NR;
  for (int t = 0; t < NT; ++t) {
    T.D[t] = sp.D[t];
    T.rate[t] = std::sqrt(2.0f * sp.D[t] * sp.dt);  // molSetTimestep (particleSystem.cpp:137-142)
  }
  OrderCsr csr[3];
  build_csr(sp, csr);

  // Per-reaction working copies of struct Reaction's fields.
  std::vector<float> prob(NR), bindR2(NR, 0.0f), unbind(NR, 0.0f), gem(NR), ppos(6 * NR, 0.0f);
  std::vector<int> code(NR, 0);
  for (int r = 0; r < NR; ++r) {
    prob[r] = sp.rxn[r].order == 2 ? (sp.rxn[r].prob > 0 ? sp.rxn[r].prob : 1.0f) : 1.0f;
    gem[r] = sp.rxn[r].pgem > 0 ? sp.rxn[r].pgem : 0.2f;
  }
  // rxnsetrate (smolreact.cpp:283-397), all reactions first so that
  // rxnsetproduct always sees the binding radius of a reverse reaction.
  for (int r = 0; r < NR; ++r) {
    const smd_reaction& R = sp.rxn[r];
    if (R.rate < 0) { code[r] = 1; continue; }
    if (R.order == 0) {
      // getSystemVolume(): the spec's volume, or the box volume when unset
      const float vol = sp.volume > 0 ? sp.volume
                                      : (sp.box_hi[0] - sp.box_lo[0]) * (sp.box_hi[1] - sp.box_lo[1]) *
                                            (sp.box_hi[2] - sp.box_lo[2]);
      prob[r] = R.rate * sp.dt * vol;
    } else if (R.order == 1) {
      const int i = R.reactants[0];
      const OrderCsr& c = csr[1];
      float sum = 0;
      for (int j = 0; j < c.count[i]; ++j) {
        const int r2 = c.list[c.start[i] + j];
        if (sp.rxn[r2].rate > 0) sum += sp.rxn[r2].rate;
      }
      if (R.rate == 0) {
        prob[r] = 0;
      } else {
        if (sum > 0) prob[r] = R.rate / sum * (1.0f - std::exp(-sp.dt * sum));
        float prior = 0;
        for (int j = 0; j < c.count[i]; ++j) {
          const int r2 = c.list[c.start[i] + j];
          if (r2 == r) break;
          prior += prob[r2];
        }
        prob[r] = prob[r] / (1.0f - prior);
      }
    } else {
      const int i1 = R.reactants[0], i2 = R.reactants[1];
      float rate3 = R.rate;
      if (i1 == i2) rate3 *= 2;
      const float dsum = sp.D[i1] + sp.D[i2];
      if (find_reverse(sp, csr, r) < 0) unbind[r] = 0.0f;
      if (rate3 <= 0) {
        bindR2[r] = 0;
      } else if (dsum <= 0) {
        code[r] = 4;
      } else {
        const float br = (float)bindingradius(rate3, sp.dt, dsum, -1, 0);
        bindR2[r] = br * br;
      }
    }
  }
  // rxnsetproduct (smolreact.cpp:410-483)
  for (int r = 0; r < NR; ++r) {
    const smd_reaction& R = sp.rxn[r];
    if (R.nprod != 2) continue;
    float dc1 = sp.D[R.products[0]], dc2 = sp.D[R.products[1]];
    float dsum = dc1 + dc2;
    const int rev = find_reverse(sp, csr, r);
    if (rev < 0) {
      if (dsum == 0) {
        dc1 = 1.0f;
        dc2 = 1.0f;
        dsum = 0.5f;
      }
      const float rpar = unbind[r];
      ppos[6 * r] = rpar * dc1 / dsum;
      ppos[6 * r + 3] = -rpar * dc2 / dsum;
    } else {
      const float bindradr = bindR2[rev] >= 0 ? std::sqrt(bindR2[rev]) : -1.0f;
      if (dsum <= 0) code[r] = 4;
      const float rpar = (float)unbindingradius(gem[r], sp.dt, dsum, bindradr);
      if (rpar == -2) {
        code[r] = 7;
      } else if (rpar < 0) {
        // The requested geminate probability exceeds the largest reachable
        // one (-rpar): products go to the reaction point, which gives that
        // largest probability.
        code[r] = 8;
        unbind[r] = 0.0f;
      } else {
        unbind[r] = rpar;
        ppos[6 * r] = rpar * dc1 / dsum;
        ppos[6 * r + 3] = -rpar * dc2 / dsum;
      }
    }
  }
  // Device tables.
  for (int t = 0; t < NT; ++t) {
    T.uniCount[t] = csr[1].count[t];
    T.uniStart[t] = csr[1].start[t];
  }
  for (size_t k = 0; k < csr[1].list.size(); ++k) T.uniList[k] = csr[1].list[k];
  for (int k = 0; k < kMaxTypes * kMaxTypes; ++k) T.pairRxn[k] = -1;
  for (int r = 0; r < NR; ++r) {
    const smd_reaction& R = sp.rxn[r];
    T.prob[r] = prob[r];
    T.bindR2[r] = bindR2[r];
    T.nprod[r] = R.nprod;
    T.prodX[r] = prod_x(R);
    T.prodY[r] = prod_y(R);
    for (int k = 0; k < 6; ++k) T.prodPos[r][k] = ppos[6 * r + k];
    if (R.order == 1) T.hasUni = 1;
    if (R.order == 2) {
      T.hasBim = 1;
      T.pairRxn[R.reactants[0] * kMaxTypes + R.reactants[1]] = r;  // last writer wins (particleSystem.cpp:656-670)
      T.pairRxn[R.reactants[1] * kMaxTypes + R.reactants[0]] = r;
    }
    if (der) {
      der[r].prob = prob[r];
      der[r].bind_radius2 = bindR2[r];
      der[r].bind_radius = std::sqrt(bindR2[r]);
      der[r].unbind_radius = unbind[r];
      for (int k = 0; k < 6; ++k) der[r].product_pos[k] = ppos[6 * r + k];
      der[r].setup_code = code[r];
    }
  }
  for (int r = 0; r < NR; ++r)
    if (code[r] != 0 && code[r] != 8) {
      err = "reaction " + std::to_string(r) + ": setup error code " + std::to_string(code[r]);
      return false;
    }

  // Order-0 creation, in reaction index order (reaction_cpu.cpp:zeroReact).
  // poisrandF (stats.cpp:234-275) uses the multiplication method below a mean
  // of 12; larger means are split into equal pieces below 12 (a sum of
  // independent Poisson variates is Poisson with the summed mean).
  double zeroMean = 0.0;
  for (int r = 0; r < NR; ++r) {
    const smd_reaction& R = sp.rxn[r];
    if (R.order != 0 || R.nprod < 1 || !(prob[r] > 0.0f)) continue;
    const int k = T.numZero++;
    const int pieces = prob[r] < 12.0f ? 1 : (int)std::ceil(prob[r] / 8.0f);
    const float xm = pieces == 1 ? prob[r] : prob[r] / (float)pieces;
    T.zeroList[k] = r;
    T.zeroPieces[k] = pieces;
    T.zeroG[k] = (float)std::exp(-(double)xm);
    zeroMean += prob[r];
  }
  // Creations beyond mean + 10 sd (+32) in one step are reported as an error.
  T.zeroMax = T.numZero ? (uint32_t)std::ceil(zeroMean + 10.0 * std::sqrt(zeroMean) + 32.0) : 0u;

  // Pair-search filters derived from the final pair table.
  for (int a = 0; a < NT; ++a) {
    T.partnerMask[a] = 0;
    T.maxBindR2[a] = 0.0f;
    for (int b = 0; b < NT; ++b) {
      const int r = T.pairRxn[a * kMaxTypes + b];
      if (r < 0) continue;
      T.partnerMask[a] |= 1u << b;
      T.maxBindR2[a] = std::fmax(T.maxBindR2[a], T.bindR2[r]);
    }
  }

  // Geometry.
  for (int d = 0; d < 3; ++d) {
    T.lo[d] = sp.box_lo[d];
    T.hi[d] = sp.box_hi[d];
    T.L[d] = sp.box_hi[d] - sp.box_lo[d];
    T.halfL[d] = 0.5f * T.L[d];
    if (!(T.L[d] > 0)) { err = "box must have positive size"; return false; }
    const bool pl = sp.face[2 * d] == kPeriodic, ph = sp.face[2 * d + 1] == kPeriodic;
    if (pl != ph) { err = "periodic faces must come in pairs"; return false; }
    T.periodic[d] = pl ? 1 : 0;
  }
  for (int f = 0; f < 6; ++f) {
    if (sp.face[f] < 0 || sp.face[f] > 3) { err = "bad face kind"; return false; }
    T.face[f] = sp.face[f];
  }
  T.particleRadius = sp.particle_radius;
  if (sp.num_spheres < 0 || sp.num_spheres > kMaxSpheres) { err = "too many spheres"; return false; }
  T.numSpheres = sp.num_spheres;
  for (int k = 0; k < sp.num_spheres; ++k)
    for (int q = 0; q < 4; ++q) T.sphere[k][q] = sp.sphere[k][q];
  T.unbindOffsets = sp.unbind_offsets ? 1 : 0;

  // Hash grid.
  float maxBind = 0;
  for (int r = 0; r < NR; ++r)
    if (sp.rxn[r].order == 2) maxBind = std::fmax(maxBind, std::sqrt(bindR2[r]));
  uint32_t totalBits = 0;
  for (int d = 0; d < 3; ++d) {
    if (sp.grid_bits[d] < 2 || sp.grid_bits[d] > 12) { err = "grid_bits must be in [2, 12]"; return false; }
    T.gbits[d] = sp.grid_bits[d];
    totalBits += sp.grid_bits[d];
    const float dim = (float)(1u << sp.grid_bits[d]);
    float cs = sp.cell_size[d] > 0 ? sp.cell_size[d] : T.L[d] / dim;
    if (T.periodic[d]) {
      // The wrapped grid must tile the periodic box exactly.
      double k = std::floor((double)T.L[d] / ((double)cs * dim) + 1e-6);
      if (k < 1) k = 1;
      cs = (float)((double)T.L[d] / (k * dim));
      if (maxBind > T.halfL[d]) { err = "binding radius exceeds half the periodic box"; return false; }
    }
    if (cs < maxBind) { err = "cell size smaller than the largest binding radius"; return false; }
    T.cellSize[d] = cs;
    T.origin[d] = T.periodic[d] ? sp.box_lo[d] : sp.origin[d];
  }
  if (totalBits > 28) { err = "grid too large"; return false; }
  T.numBuckets = 1u << totalBits;
  T.numStreams = sp.num_streams ? sp.num_streams : (uint32_t)kMtParamSets;
  return true;
}

// ---------------------------------------------------------------------------
// MersenneTwister.dat (32768 records of matrix_a, mask_b, mask_c, seed).
// ---------------------------------------------------------------------------
namespace {
std::mutex g_datMutex;
std::string g_datPath;
std::vector<uint32_t> g_dat;
}  // namespace

void set_mt_dat_path(const char* p) {
  std::lock_guard<std::mutex> lk(g_datMutex);
  g_datPath = p ? p : "";
  g_dat.clear();
}

const uint32_t* mt_dat(std::string& err) {
  std::lock_guard<std::mutex> lk(g_datMutex);
  if (!g_dat.empty()) return g_dat.data();
  std::string path = g_datPath;
  if (path.empty()) {
    const char* env = std::getenv("SMD_MT_DAT");
    if (env && *env) path = env;
  }
  if (path.empty()) path = default_data_dir() + "/MersenneTwister.dat";
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) {
    err = "cannot open " + path;
    return nullptr;
  }
  std::vector<uint32_t> buf(4 * (size_t)kMtParamSets);
  const size_t got = std::fread(buf.data(), sizeof(uint32_t), buf.size(), f);
  std::fclose(f);
  if (got != buf.size()) {
    err = "short read of " + path;
    return nullptr;
  }
  g_dat.swap(buf);
  return g_dat.data();
}

// dcmt generator on the host (same recurrence as the device Mt, block form
// of rnd/genmtrand.cpp:52-87).
HostMt::HostMt(uint32_t a_, uint32_t mb_, uint32_t mc_, uint32_t seed) : a(a_), mb(mb_), mc(mc_) {
  st[0] = seed;
  for (int i = 1; i < kMtNN; ++i) st[i] = 1812433253u * (st[i - 1] ^ (st[i - 1] >> 30)) + (uint32_t)i;
  idx = 0;
}

uint32_t HostMt::next() {
  const int i1 = idx + 1 >= kMtNN ? idx + 1 - kMtNN : idx + 1;
  const int iM = idx + kMtMM >= kMtNN ? idx + kMtMM - kMtNN : idx + kMtMM;
  uint32_t x = (st[idx] & 0xFFFFFFFEu) | (st[i1] & 1u);
  x = st[iM] ^ (x >> 1) ^ ((x & 1u) ? a : 0u);
  st[idx] = x;
  idx = i1;
  x ^= x >> 12;
  x ^= (x << 7) & mb;
  x ^= (x << 15) & mc;
  x ^= x >> 18;
  return x;
}

bool stream_seeds(uint32_t seed, uint32_t count, uint32_t* out, std::string& err) {
  const uint32_t* dat = mt_dat(err);
  if (!dat) return false;
  HostMt g(dat[0], dat[1], dat[2], seed);
  for (uint32_t k = 0; k < count; ++k) out[k] = g.next();
  return true;
}

}  // namespace smd
