// Thin C-linkage shim over the reference's own CPU sources, compiled only by
// oracle/Makefile (target `ref`) into oracle/_ref/libsmdref.so.  It is test
// infrastructure: it exists so tests/golden/make_fixtures.py can record the
// reference's outputs as golden vectors.  Nothing here is shipped or measured.
//
// Wrapped reference functions:
//   sgenrand_mt / genrand_mt     rnd/genmtrand.cpp:38-87   (dcmt CPU generator)
//   createGaussTable             stats.cpp:277-282
//   numrxnrate / actrxnrate      rxnparam.cpp:42-239
//   bindingradius                rxnparam.cpp:253-277
//   unbindingradius              rxnparam.cpp:287-308
//   rdfsteadystate               rxnparam.cpp:451-475
//   init_gen_rand / unirandOOF   SFMT/SFMT.cpp, stats.h:35-37 (randPos)
#include <cstddef>
#include <cstdint>

#include "rnd/dci.h"
#include "rxnparam.h"
#include "stats.h"

void createGaussTable(float* gaussianLookupTable, size_t gaussianTableDim);

extern "C" {

// Runs the dcmt generator configured for MEXP 607 (nn=19, mm=9, r=1, w=32,
// tempering shifts 12/7/15/18) with one MersenneTwister.dat record's
// matrix_a/mask_b/mask_c, seeded with `seed`, and writes `count` outputs.
void smdref_mt_stream(uint32_t matrix_a, uint32_t mask_b, uint32_t mask_c,
                      uint32_t seed, int count, uint32_t* out) {
  uint32_t st[19];
  mt_struct m;
  m.aaa = matrix_a;
  m.mm = 9;
  m.nn = 19;
  m.rr = 1;
  m.ww = 32;
  m.wmask = 0xFFFFFFFFu;
  m.umask = 0xFFFFFFFEu;
  m.lmask = 0x1u;
  m.shift0 = 12;
  m.shift1 = 18;
  m.shiftB = 7;
  m.shiftC = 15;
  m.maskB = mask_b;
  m.maskC = mask_c;
  m.i = 0;
  m.state = st;
  sgenrand_mt(seed, &m);
  for (int k = 0; k < count; ++k) out[k] = genrand_mt(&m);
}

void smdref_gauss_table(float* out, int n) { createGaussTable(out, (size_t)n); }

double smdref_numrxnrate(double step, double a, double b) { return numrxnrate(step, a, b); }
double smdref_actrxnrate(double step, double a) { return actrxnrate(step, a); }
double smdref_bindingradius(double rate, double dt, double difc, double b, int rel) {
  return bindingradius(rate, dt, difc, b, rel);
}
double smdref_unbindingradius(double pgem, double dt, double difc, double a) {
  return unbindingradius(pgem, dt, difc, a);
}

void smdref_sfmt_uniform(uint32_t seed, int count, float lo, float hi, float* out) {
  init_gen_rand(seed);
  for (int k = 0; k < count; ++k) out[k] = unirandOOF(lo, hi);
}

double smdref_rdf_steady_state(double* r, double* rdfa, double* rdfd, int n, double step, double b, double eps) {
  return rdfsteadystate(r, rdfa, rdfd, n, step, b, eps);
}

}  // extern "C"
