// C-ABI of the simulation API (include/gpusmoldyn.h).
#include <dlfcn.h>
#include <algorithm>
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "smd_engine.h"

namespace smd {

static thread_local std::string g_lastError;

void set_error(const std::string& msg) { g_lastError = msg; }

std::string default_data_dir() {
  Dl_info info;
  if (dladdr((void*)&default_data_dir, &info) && info.dli_fname) {
    std::string p(info.dli_fname);
    const size_t slash = p.rfind('/');
    if (slash != std::string::npos) return p.substr(0, slash) + "/data";
  }
  return "data";
}

}  // namespace smd

using smd::Engine;
using smd::set_error;

#define CAPI_HIPCHK(x)                                                              \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) {                                                         \
      set_error(std::string("HIP error '") + hipGetErrorString(e_) + "' in " #x);   \
      return SMD_ERR_HIP;                                                           \
    }                                                                               \
  } while (0)

extern "C" {

const char* smd_last_error(void) { return smd::g_lastError.c_str(); }

int smd_version(void) { return 1; }

smd_status smd_set_mt_dat_path(const char* path) {
  smd::set_mt_dat_path(path);
  return SMD_OK;
}

smd_status smd_gauss_table(float* out, int n) {
  if (!out || n <= 0 || (n & 1)) {
    set_error("gauss table size must be positive and even");
    return SMD_ERR_INVALID;
  }
  smd::gauss_table(out, n);
  return SMD_OK;
}

double smd_numrxnrate(double step, double a, double b) { return smd::numrxnrate(step, a, b); }
double smd_actrxnrate(double step, double a) { return smd::actrxnrate(step, a); }
double smd_bindingradius(double rate, double dt, double difc, double b, int rel) {
  return smd::bindingradius(rate, dt, difc, b, rel);
}
double smd_unbindingradius(double pgem, double dt, double difc, double a) {
  return smd::unbindingradius(pgem, dt, difc, a);
}

double smd_rdf_steady_state(const double* r, double* rdfa, double* rdfd, int n, double step, double b, double eps) {
  if (!r || !rdfa || !rdfd || n < 10 || !(step > 0) || !(eps > 0)) return -1.0;
  return smd::rdf_steady_state(r, rdfa, rdfd, n, step, b, eps);
}

smd_status smd_rdf_sweep(char mode, int n, double eps, int down, double b, double* flux31) {
  if (!flux31 || n < 10 || !(eps > 0) || (mode != 'i' && mode != 'r' && mode != 'b') ||
      (mode == 'r' && !(b >= 1.0)) || (mode == 'b' && !(b >= 0.0 && b <= 1.0))) {
    set_error("invalid rdf sweep arguments");
    return SMD_ERR_INVALID;
  }
  return smd::rdf_sweep(mode, n, eps, down != 0, b, flux31) == 0 ? SMD_OK : SMD_ERR_INVALID;
}

smd_status smd_rdf_grid(double rmax, int n, double* r) {
  if (!r || n < 10 || !(rmax > 1.1)) {
    set_error("invalid rdf grid arguments");
    return SMD_ERR_INVALID;
  }
  std::vector<double> tmp(n + 2);
  smd::rdf_grid(rmax, n, tmp.data());
  std::memcpy(r, tmp.data(), n * sizeof(double));
  return SMD_OK;
}

smd_status smd_spec_derive(const smd_sim_spec* spec, smd_rxn_derived* out) {
  if (!spec) {
    set_error("null spec");
    return SMD_ERR_INVALID;
  }
  smd::ReplicaTables T;
  std::string err;
  std::vector<smd_rxn_derived> der(spec->num_reactions > 0 ? spec->num_reactions : 1);
  const bool ok = smd::derive_tables(*spec, T, der.data(), err);
  if (out && spec->num_reactions > 0) std::memcpy(out, der.data(), sizeof(smd_rxn_derived) * spec->num_reactions);
  if (!ok) {
    set_error(err);
    return SMD_ERR_RXN_SETUP;
  }
  return SMD_OK;
}

smd_status smd_spec_cells(const smd_sim_spec* spec, float* cell3, float* origin3) {
  if (!spec) return SMD_ERR_INVALID;
  smd::ReplicaTables T;
  std::string err;
  std::vector<smd_rxn_derived> der(spec->num_reactions > 0 ? spec->num_reactions : 1);
  if (!smd::derive_tables(*spec, T, der.data(), err)) {
    set_error(err);
    return SMD_ERR_RXN_SETUP;
  }
  for (int d = 0; d < 3; ++d) {
    if (cell3) cell3[d] = T.cellSize[d];
    if (origin3) origin3[d] = T.origin[d];
  }
  return SMD_OK;
}

smd_status smd_stream_seeds(unsigned seed, unsigned count, uint32_t* out) {
  std::string err;
  if (!smd::stream_seeds(seed, count, out, err)) {
    set_error(err);
    return SMD_ERR_IO;
  }
  return SMD_OK;
}

smd_status smd_sfmt_uniform(unsigned seed, unsigned count, float lo, float hi, float* out) {
  if (count && !out) {
    set_error("null output");
    return SMD_ERR_INVALID;
  }
  smd::Sfmt g(seed);
  for (unsigned k = 0; k < count; ++k) out[k] = g.uniform_open(lo, hi);
  return SMD_OK;
}

smd_status smd_mt_params(unsigned index, uint32_t* abc) {
  std::string err;
  const uint32_t* dat = smd::mt_dat(err);
  if (!dat) {
    set_error(err);
    return SMD_ERR_IO;
  }
  if (index >= (unsigned)smd::kMtParamSets) {
    set_error("record index out of range");
    return SMD_ERR_INVALID;
  }
  abc[0] = dat[4 * index];
  abc[1] = dat[4 * index + 1];
  abc[2] = dat[4 * index + 2];
  return SMD_OK;
}

smd_status smd_device_count(int* n) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  *n = c;
  return SMD_OK;
}

smd_engine* smd_engine_create(int device, int radix_bits) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess || c <= 0) {
    set_error("no HIP device visible");
    return nullptr;
  }
  if (device < 0 || device >= c) {
    set_error("device index out of range");
    return nullptr;
  }
  if (radix_bits != 0 && radix_bits != 4 && radix_bits != 8 && radix_bits != 11) {
    set_error("radix_bits must be 0, 4, 8 or 11");
    return nullptr;
  }
  smd_engine* e = new smd_engine;
  e->impl = new Engine(device, radix_bits);
  return e;
}

void smd_engine_destroy(smd_engine* e) {
  if (!e) return;
  delete e->impl;
  delete e;
}

int smd_engine_add_replica(smd_engine* e, const smd_sim_spec* spec, const float* pos4, const int* types,
                           unsigned n) {
  if (!e || !spec || (n && (!pos4 || !types))) {
    set_error("null argument");
    return -SMD_ERR_INVALID;
  }
  return e->impl->add_replica(*spec, pos4, types, n);
}

smd_status smd_engine_finalize(smd_engine* e, unsigned capacity) {
  if (!e) return SMD_ERR_INVALID;
  return e->impl->finalize(capacity);
}

smd_status smd_engine_finalize_slots(smd_engine* e, unsigned n_slots, unsigned slot_bucket_bits,
                                     unsigned slot_streams, unsigned capacity, unsigned inject_cap) {
  if (!e) return SMD_ERR_INVALID;
  return e->impl->finalize_slots(n_slots, slot_bucket_bits, slot_streams, capacity, inject_cap);
}

smd_status smd_engine_admit(smd_engine* e, int slot, const smd_sim_spec* spec, const float* pos4, const int* types,
                            unsigned n) {
  if (!e || !spec || (n && (!pos4 || !types))) return SMD_ERR_INVALID;
  return e->impl->admit(slot, *spec, pos4, types, n);
}

smd_status smd_engine_retire(smd_engine* e, int slot) {
  if (!e) return SMD_ERR_INVALID;
  return e->impl->retire(slot);
}

smd_status smd_engine_step(smd_engine* e, int nsteps) {
  if (!e || nsteps < 0) return SMD_ERR_INVALID;
  return e->impl->step(nsteps);
}

smd_status smd_engine_sync(smd_engine* e) {
  if (!e) return SMD_ERR_INVALID;
  return e->impl->sync();
}

smd_status smd_engine_counts(smd_engine* e, unsigned* counts) {
  if (!e || !counts) return SMD_ERR_INVALID;
  return e->impl->counts(counts);
}

smd_status smd_engine_set_graphs(smd_engine* e, int on) {
  if (!e) return SMD_ERR_INVALID;
  e->impl->useGraphs = on != 0;
  if (!on) e->impl->drop_graph();
  return SMD_OK;
}

smd_status smd_engine_set_knob(smd_engine* e, int knob, int value) {
  if (!e) return SMD_ERR_INVALID;
  return e->impl->set_knob(knob, value);
}

int smd_engine_get_knob(const smd_engine* e, int knob) {
  if (!e) return -SMD_ERR_INVALID;
  if (knob == SMD_KNOB_PRODUCER_BLOCK) return e->impl->producerBlock;
  if (knob == SMD_KNOB_PAIR_BLOCK) return e->impl->pairBlock;
  if (knob == SMD_KNOB_SEGMENTED_SORT) return e->impl->segSort;
  return -SMD_ERR_INVALID;
}

smd_status smd_engine_set_sort(smd_engine* e, int kind) {
  if (!e || (kind != 0 && kind != 1)) return SMD_ERR_INVALID;
  e->impl->drop_graph();
  e->impl->sb.algo = kind;
  return SMD_OK;
}

smd_status smd_engine_particle_steps(smd_engine* e, unsigned long long* out) {
  if (!e || !out) return SMD_ERR_INVALID;
  return e->impl->particle_steps(out);
}

int smd_engine_radix_bits(const smd_engine* e) { return e ? e->impl->radixBits : 0; }

unsigned smd_engine_num_replicas(const smd_engine* e) { return e ? (unsigned)e->impl->tabs.size() : 0u; }

smd_status smd_engine_get(smd_engine* e, int r, float* pos4, int* types, unsigned cap, unsigned* n_out) {
  if (!e) return SMD_ERR_INVALID;
  return e->impl->get(r, pos4, types, cap, n_out);
}

smd_status smd_engine_species_counts(smd_engine* e, int r, unsigned* counts) {
  if (!e || !counts) return SMD_ERR_INVALID;
  unsigned n = 0;
  smd_status st = e->impl->get(r, nullptr, nullptr, 0, &n);
  if (st != SMD_OK) return st;
  std::vector<int> t(n > 0 ? n : 1);
  st = e->impl->get(r, nullptr, t.data(), n, &n);
  if (st != SMD_OK) return st;
  const int NT = e->impl->specs[r].num_types;
  for (int k = 0; k < NT; ++k) counts[k] = 0;
  for (unsigned i = 0; i < n; ++i)
    if (t[i] >= 0 && t[i] < NT) counts[t[i]]++;
  return SMD_OK;
}

smd_status smd_engine_get_streams(smd_engine* e, int r, uint32_t* words, unsigned cap_streams) {
  if (!e || !words) return SMD_ERR_INVALID;
  return e->impl->get_streams(r, words, cap_streams, false);
}

smd_status smd_engine_set_streams(smd_engine* e, int r, const uint32_t* words, unsigned n_streams) {
  if (!e || !words) return SMD_ERR_INVALID;
  return e->impl->get_streams(r, const_cast<uint32_t*>(words), n_streams, true);
}

smd_status smd_engine_rxn(smd_engine* e, int r, smd_rxn_derived* out) {
  if (!e || !out || r < 0 || (size_t)r >= e->impl->derived.size()) return SMD_ERR_INVALID;
  const int nr = e->impl->specs[r].num_reactions;
  if (nr > 0) std::memcpy(out, e->impl->derived[r].data(), sizeof(smd_rxn_derived) * nr);
  return SMD_OK;
}

smd_status smd_engine_debug_stage(smd_engine* e, int stage, float* pos4, int* types, uint32_t* keys_sorted,
                                  uint32_t* vals_sorted, int* cand, uint32_t* n_total, uint32_t* n_live) {
  if (!e || stage < 1 || stage > 4) return SMD_ERR_INVALID;
  return e->impl->debug_stage(stage, pos4, types, keys_sorted, vals_sorted, cand, n_total, n_live);
}

smd_status smd_engine_debug_first(smd_engine* e, uint32_t* first, uint32_t cap, uint32_t* n_out) {
  if (!e || !e->impl->finalized) return SMD_ERR_INVALID;
  const uint32_t words = e->impl->nb + 1;
  if (n_out) *n_out = words;
  if (first && cap) {
    CAPI_HIPCHK(hipSetDevice(e->impl->device));
    CAPI_HIPCHK(hipStreamSynchronize(e->impl->stream));
    CAPI_HIPCHK(hipMemcpy(first, e->impl->dFirst, std::min(cap, words) * sizeof(uint32_t), hipMemcpyDeviceToHost));
  }
  return SMD_OK;
}

smd_status smd_engine_layout(smd_engine* e, uint32_t* total_buckets, uint32_t* key_bits, uint32_t* total_streams,
                             uint32_t* capacity) {
  if (!e || !e->impl->finalized) return SMD_ERR_INVALID;
  if (total_buckets) *total_buckets = e->impl->deadKey;
  if (key_bits) *key_bits = e->impl->keyBits;
  if (total_streams) *total_streams = e->impl->Stot;
  if (capacity) *capacity = e->impl->cap;
  return SMD_OK;
}

smd_status smd_engine_sort_plan(const smd_engine* e, int* passes, int* segmented) {
  if (!e || !e->impl->finalized) return SMD_ERR_INVALID;
  const bool seg = e->impl->seg_sorted();
  if (segmented) *segmented = seg ? 1 : 0;
  if (passes) *passes = seg ? 2 : smd::sort_passes(e->impl->keyBits, e->impl->radixBits);
  return SMD_OK;
}

smd_status smd_engine_profile_step(smd_engine* e, int nsteps, double* us, int max_stages, char* names, int names_cap,
                                   int* n_stages) {
  if (!e || !us) return SMD_ERR_INVALID;
  return e->impl->profile(nsteps, us, max_stages, names, names_cap, n_stages);
}

smd_status smd_mt_generate(int device, const uint32_t* seeds, unsigned n_streams, int count, int batch3,
                           uint32_t* out) {
  if (!seeds || !out || count < 0) return SMD_ERR_INVALID;
  std::string err;
  const uint32_t* dat = smd::mt_dat(err);
  if (!dat) {
    set_error(err);
    return SMD_ERR_IO;
  }
  CAPI_HIPCHK(hipSetDevice(device));
  const uint32_t S = n_streams;
  std::vector<uint32_t> par(3 * (size_t)S);
  for (uint32_t k = 0; k < S; ++k) {
    const uint32_t rec = k % smd::kMtParamSets;
    par[k] = dat[4 * rec];
    par[(size_t)S + k] = dat[4 * rec + 1];
    par[2 * (size_t)S + k] = dat[4 * rec + 2];
  }
  uint32_t *dSt = nullptr, *dPar = nullptr, *dSeeds = nullptr, *dOut = nullptr;
  CAPI_HIPCHK(hipMalloc((void**)&dSt, (size_t)smd::kMtWords * S * 4 + 4));
  CAPI_HIPCHK(hipMalloc((void**)&dPar, par.size() * 4 + 4));
  CAPI_HIPCHK(hipMalloc((void**)&dSeeds, (size_t)S * 4 + 4));
  CAPI_HIPCHK(hipMalloc((void**)&dOut, (size_t)S * count * 4 + 4));
  CAPI_HIPCHK(hipMemcpy(dPar, par.data(), par.size() * 4, hipMemcpyHostToDevice));
  CAPI_HIPCHK(hipMemcpy(dSeeds, seeds, (size_t)S * 4, hipMemcpyHostToDevice));
  CAPI_HIPCHK(smd::launch_mt_init(dSt, dSeeds, S, nullptr));
  CAPI_HIPCHK(smd::launch_mt_fill(dSt, dPar, S, S, count, batch3, dOut, nullptr));
  CAPI_HIPCHK(hipDeviceSynchronize());
  CAPI_HIPCHK(hipMemcpy(out, dOut, (size_t)S * count * 4, hipMemcpyDeviceToHost));
  (void)hipFree(dSt);
  (void)hipFree(dPar);
  (void)hipFree(dSeeds);
  (void)hipFree(dOut);
  return SMD_OK;
}

smd_status smd_sort_pairs(int device, uint32_t* keys, uint32_t* values, unsigned n, unsigned key_bits,
                          int radix_bits) {
  if (n == 0) return SMD_OK;
  if (!keys || !values || key_bits == 0 || key_bits > 32 || (radix_bits != 4 && radix_bits != 8 && radix_bits != 11)) {
    set_error("invalid sort arguments");
    return SMD_ERR_INVALID;
  }
  CAPI_HIPCHK(hipSetDevice(device));
  smd::SortBuffers b;
  b.status = nullptr;
  b.algo = 1;
  b.gatherSrc = nullptr;
  b.gatherDst = nullptr;
  b.gatherDeadKey = 0;
  b.sizeHint = 0;
  b.packIndex = 0;
  uint32_t* dN = nullptr;
  CAPI_HIPCHK(hipMalloc((void**)&b.keys, (size_t)n * 4));
  CAPI_HIPCHK(hipMalloc((void**)&b.vals, (size_t)n * 4));
  CAPI_HIPCHK(hipMalloc((void**)&b.keysAlt, (size_t)n * 4));
  CAPI_HIPCHK(hipMalloc((void**)&b.valsAlt, (size_t)n * 4));
  CAPI_HIPCHK(hipMalloc((void**)&b.counts, smd::sort_counts_words(n, radix_bits) * 4));
  CAPI_HIPCHK(hipMalloc((void**)&dN, 4));
  CAPI_HIPCHK(hipMemcpy(b.keys, keys, (size_t)n * 4, hipMemcpyHostToDevice));
  CAPI_HIPCHK(hipMemcpy(b.vals, values, (size_t)n * 4, hipMemcpyHostToDevice));
  CAPI_HIPCHK(hipMemcpy(dN, &n, 4, hipMemcpyHostToDevice));
  uint32_t* k0 = b.keys;
  uint32_t* v0 = b.vals;
  uint32_t* k1 = b.keysAlt;
  uint32_t* v1 = b.valsAlt;
  CAPI_HIPCHK(smd::launch_sort(b, dN, n, key_bits, radix_bits, false, nullptr));
  CAPI_HIPCHK(hipDeviceSynchronize());
  CAPI_HIPCHK(hipMemcpy(keys, b.keys, (size_t)n * 4, hipMemcpyDeviceToHost));
  CAPI_HIPCHK(hipMemcpy(values, b.vals, (size_t)n * 4, hipMemcpyDeviceToHost));
  (void)hipFree(k0);
  (void)hipFree(v0);
  (void)hipFree(k1);
  (void)hipFree(v1);
  (void)hipFree(b.counts);
  (void)hipFree(dN);
  return SMD_OK;
}

}  // extern "C"
