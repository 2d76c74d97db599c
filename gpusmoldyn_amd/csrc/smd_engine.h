// The ensemble engine behind the smd_engine C handle.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "smd_kernels.h"
#include "smd_setup.h"

namespace smd {

struct StageTimer {
  hipEvent_t ev[8];
};

class Engine {
 public:
  Engine(int device, int radixBits);
  ~Engine();
  int add_replica(const smd_sim_spec& spec, const float* pos4, const int* types, uint32_t n);
  smd_status finalize(uint32_t capacity);
  smd_status step(int nsteps);
  smd_status sync();
  smd_status counts(uint32_t* out);
  smd_status get(int r, float* pos4, int* types, uint32_t capN, uint32_t* nOut);
  smd_status get_streams(int r, uint32_t* words, uint32_t capStreams, bool set);
  smd_status debug_stage(int stage, float* pos4, int* types, uint32_t* keysSorted, uint32_t* valsSorted, int* cand,
                         uint32_t* nTotal, uint32_t* nLive);
  smd_status profile(int nsteps, double* us, int maxStages, char* names, int namesCap, int* nStages);
  smd_status particle_steps(unsigned long long* out);
  // Replace the particles of a single-replica engine (setArray / reset).
  smd_status set_particles(const float* pos4, const int* types, uint32_t n);
  // Device arrays of the current particles in the caller layout (pos4 =
  // (x, y, z, w), int types), unpacked from the canonical state into view
  // buffers that stay valid until the next step or view call.
  smd_status view(float4** pos4, int** types, uint32_t* n);

  int device;
  int radixBits;
  bool finalized = false;
  bool broken = false;
  hipStream_t stream = nullptr;
  std::vector<smd_sim_spec> specs;
  std::vector<ReplicaTables> tabs;
  std::vector<std::vector<smd_rxn_derived>> derived;
  std::vector<std::vector<float>> pos0;
  std::vector<std::vector<int>> type0;
  uint32_t R = 0, Stot = 0, deadKey = 0, keyBits = 0, nb = 0, cap = 0, nTilesCap = 0;
  bool anySplit = false, anyBim = false, anyZero = false;
  uint32_t zeroTotal = 0;  // order-0 creations per step, all replicas, at most
  uint64_t Nub = 0, knownN = 0, stepsDone = 0;
  int cur = 0;

  // device state
  ReplicaTables* dTables = nullptr;
  int* dStreamReplica = nullptr;
  uint32_t* dBucketStart = nullptr;
  int* dBucketRep = nullptr;  // replica of each 2^repShift-bucket chunk (nullptr: search bucketStart)
  uint32_t repShift = 0;
  uint32_t *dRepOff = nullptr, *dRepCount = nullptr, *dNewOff = nullptr, *dNewCount = nullptr;
  uint32_t *dAliveBase = nullptr, *dBirthBase = nullptr, *dRepAlive = nullptr;
  uint32_t *dRepDeaths = nullptr, *dRepBirths = nullptr;
  uint32_t* dCounters = nullptr;
  unsigned long long* dParticleSteps = nullptr;  // running sum of particles entering each step
  uint32_t *dN = nullptr, *dNnext = nullptr, *dLive = nullptr;
  uint32_t *dMtState = nullptr, *dMtPar = nullptr;
  float* dGauss = nullptr;
  uint32_t* dFirst = nullptr;
  // canonical particle state, ping-ponged by the compaction: (x, y, z, type
  // bits) and the fourth position component
  float4* dPT[2] = {nullptr, nullptr};
  float* dW[2] = {nullptr, nullptr};
  float4* dAddedPos = nullptr;
  int* dBirthType = nullptr;
  float4* dSortedPT = nullptr;
  int* dCand = nullptr;
  float4* dViewPos = nullptr;  // view() buffers
  int* dViewType = nullptr;
  uint32_t viewCap = 0;
  SortBuffers sb = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 1, nullptr, nullptr, 0, 0};
  uint32_t *dTileAlive = nullptr, *dTileBirth = nullptr, *dTileAlivePre = nullptr, *dTileBirthPre = nullptr;
  uint32_t* dBlockSum = nullptr;  // finalize block totals, [2][blockStride]
  uint32_t blockStride = 0;
  uint32_t* dStatus = nullptr;  // sticky kernel error bits
  // order-0 creation: one stream per replica, a per-step creation buffer
  uint32_t *dZMt = nullptr, *dZPar = nullptr, *dZCount = nullptr, *dZeroBase = nullptr;
  float4* dZPos = nullptr;
  int* dZType = nullptr;

  // particle count after each step, written by the finalize kernel into
  // host-visible memory: slot (step & (kRing - 1)) = step << 32 | count
  static constexpr uint32_t kRing = 64;
  unsigned long long* countRing = nullptr;   // host pointer
  unsigned long long* dCountRing = nullptr;  // its device address
  uint32_t* dStepCount = nullptr;
  uint64_t lastSeen = 0;  // latest step whose count has been read

  // Serving: fixed replica slots filled and emptied between steps.
  smd_status finalize_slots(uint32_t nSlots, uint32_t slotBucketBits, uint32_t slotStreams, uint32_t capacity,
                            uint32_t injectCap);
  smd_status admit(int slot, const smd_sim_spec& spec, const float* pos4, const int* types, uint32_t n);
  smd_status retire(int slot);
  bool slotsMode = false;
  uint32_t slotBits = 0, slotStreams = 0;
  uint32_t *dRetire = nullptr, *dInjCount = nullptr, *dInjBase = nullptr, *dInjSrcOff = nullptr;
  float4* dInjPos = nullptr;
  int* dInjType = nullptr;
  int* dInjRep = nullptr;
  uint32_t injCap = 0, injN = 0;
  std::vector<uint8_t> slotBusy;     // admitted and not yet retired
  std::vector<uint8_t> slotPending;  // admitted, particles not yet entered

  // launch shapes and the sort's form (identical results for every setting;
  // see set_knob)
  int producerBlock = 256, pairBlock = 64, segSort = 1;
  int segLow = 0, segHigh = 0;       // digit widths of the segmented sort, 0: not applicable
  uint32_t* dSegScratch = nullptr;   // its tile table and counters
  smd_status set_knob(int knob, int value);

  // hipGraph replay of two-step pairs, one executable graph per ping-pong
  // parity (an eager step between pairs flips the parity)
  bool useGraphs = true;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  uint32_t gMaxN[2] = {0, 0};
  // Waits for the stream, then destroys the executable graph of parity
  // `which` (-1: both).
  void drop_graph(int which = -1);

  bool seg_sorted() const;  // this engine's step sorts replica by replica

 private:
  smd_status enqueue_kernels(uint32_t maxN, StageTimer* timer);
  SegSortArgs seg_args() const;
  hipError_t sort_keys(uint32_t maxN, bool tilesLaid = false);
  smd_status step_graph(int pairs);
  smd_status alloc_particle_buffers(uint32_t c);
  void free_particle_buffers();
  void refresh_bound(bool block);
  smd_status grow(uint64_t need);
  smd_status ensure_capacity();
  // Upper bound of the particle count after one step that starts with x.
  uint64_t step_bound(uint64_t x) const { return (anySplit ? 2 * x : x) + zeroTotal; }
  smd_status check_status();
  smd_status enqueue_step(StageTimer* timer);
  // Host-side buffer selection that enqueue_kernels advances by one step.
  struct HostSnapshot {
    int cur;
    uint32_t *repOff, *newOff, *repCount, *newCount, *n, *nnext;
    SortBuffers sb;
  };
  HostSnapshot snapshot() const;
  void restore(const HostSnapshot& s);
};

}  // namespace smd

// The opaque C handle of include/gpusmoldyn.h.
struct smd_engine {
  smd::Engine* impl;
};
