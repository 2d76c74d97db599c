// Host-side setup routines (setup.cpp) and shared host helpers.
#pragma once
#include <stdint.h>

#include <string>

#include "../../include/gpusmoldyn.h"
#include "smd_types.h"

namespace smd {

void gauss_table(float* out, int n);
double numrxnrate(double step, double a, double b);
double actrxnrate(double step, double a);
double bindingradius(double rate, double dt, double difc, double b, int rel);
double unbindingradius(double pgem, double dt, double difc, double a);
double erfc_d(double x);  // rxnparam_erfccD (rxnparam.cpp:316-322)
// rdf tables (rxnparam.cpp:325-564), reduced units (binding radius 1)
double rdf_absorb(const double* r, double* rdf, int n);
void rdf_diffuse(const double* r, const double* rdfa, double* rdfd, int n, double step);
void rdf_reverse(const double* r, double* rdf, int n, double step, double b, double flux);
double rdf_steady_state(const double* r, double* rdfa, double* rdfd, int n, double step, double b, double eps);
void rdf_grid(double rmax, int n, double* r);
int rdf_sweep(char mode, int n, double eps, bool down, double b, double* flux31);

// Builds the device tables of one replica from its spec (fillReactionOrderInfo,
// rxnsetrate, rxnsetproduct, molSetTimestep, grid and geometry checks).
// `der` (optional) receives num_reactions derived records.
bool derive_tables(const smd_sim_spec& sp, ReplicaTables& T, smd_rxn_derived* der, std::string& err);

void set_mt_dat_path(const char* p);
const uint32_t* mt_dat(std::string& err);  // 4 words per record
std::string default_data_dir();

struct HostMt {
  uint32_t st[kMtNN];
  int idx;
  uint32_t a, mb, mc;
  HostMt(uint32_t a, uint32_t mb, uint32_t mc, uint32_t seed);
  uint32_t next();
};

bool stream_seeds(uint32_t seed, uint32_t count, uint32_t* out, std::string& err);

// SFMT-19937 (sfmt.cpp): the host generator behind the reference's
// init_gen_rand / genrand_real3 / unirandOOF (stats.h:35-37).
struct Sfmt {
  static constexpr int kWords128 = 156, kWords32 = 4 * kWords128;
  uint32_t st[kWords128][4];
  int idx = kWords32;
  explicit Sfmt(uint32_t s = 0) { seed(s); }
  void seed(uint32_t s);        // init_gen_rand
  uint32_t next32();            // gen_rand32
  double real_open();           // genrand_real3: (0, 1)
  float uniform_open(float lo, float hi);  // unirandOOF
 private:
  void refill();
};

void set_error(const std::string& msg);

}  // namespace smd
