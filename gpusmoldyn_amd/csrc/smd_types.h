// Plain-old-data structures shared by the host orchestrator and the HIP
// kernels.  Everything the device reads per replica lives in ReplicaTables
// (one ~3 KB record per replica, read through the L1/L2 caches); everything
// that changes every step lives in the flat arrays of EnsembleDev.
#pragma once
#include <stdint.h>

namespace smd {

constexpr int kMaxTypes = 16;    // species + the dummy type 0
constexpr int kMaxRxn = 32;      // reactions per network
constexpr int kMaxSpheres = 4;   // reflective spheres per replica
constexpr int kMtNN = 19;        // dcmt MEXP 607 state words (MersenneTwister.h:30)
constexpr int kMtMM = 9;         // middle word offset (MersenneTwister.cuh:29)
constexpr int kMtWords = 20;     // 19 state words + iState, stored word-major (SoA)
constexpr int kMtParamSets = 32768;  // records in MersenneTwister.dat (MersenneTwister.h:31)
constexpr int kGaussDim = 4096;  // particles_kernel.cuh:43
constexpr int kCompactTile = 1024;   // elements per compaction tile (4 waves x 4 chunks of 64)
// Sorted particle records may carry the particle's index above its type:
// w bits = index << kIndexShift | type (types stay below 2^kIndexShift), for
// indices below kMaxPackedIndex.
constexpr int kIndexShift = 5;
constexpr int kTypeMask = (1 << kIndexShift) - 1;
constexpr uint32_t kMaxPackedIndex = 1u << (32 - kIndexShift);
static_assert(kMaxTypes <= (1 << kIndexShift), "types fit below the packed index");

// Boundary behaviour of one face of the simulation box.
enum Boundary : int32_t {
  kClamp = 0,     // particles_kernel.cu:60-95: clamp to [lo+r, hi-r]
  kReflect = 1,   // mirror across the face
  kAbsorb = 2,    // remove the particle
  kPeriodic = 3,  // wrap (both faces of the axis must be periodic)
};

// Sticky engine status bits, set by kernels and reported by the host.
constexpr uint32_t kStatusSortSpin = 1u;      // a sort look-back wait hit its bound
constexpr uint32_t kStatusZeroOverflow = 2u;  // order-0 creations exceeded the per-step limit
constexpr uint32_t kStatusSortLayout = 4u;    // a segmented sort destination fell outside the particles

struct ReplicaTables {
  int32_t numTypes;
  int32_t numRxn;
  int32_t hasUni;          // any order-1 reaction
  int32_t hasBim;          // any order-2 reaction
  float rate[kMaxTypes];   // rms step per axis, sqrtf(2 D dt)  (particleSystem.cpp:137-142)
  float D[kMaxTypes];      // diffusion coefficients (bimReact weighting)
  // order-1 CSR over reactant type (reaction.h:52-66, smolreact.cpp:173-271)
  int32_t uniCount[kMaxTypes];
  int32_t uniStart[kMaxTypes];
  int32_t uniList[kMaxRxn];
  // per reaction (reaction.h:36-48)
  float prob[kMaxRxn];
  float bindR2[kMaxRxn];
  int32_t nprod[kMaxRxn];
  int32_t prodX[kMaxRxn];
  int32_t prodY[kMaxRxn];
  float prodPos[kMaxRxn][6];
  // order-2 reaction id per (type, type); -1 = none (particleSystem.cpp:651-673)
  int32_t pairRxn[kMaxTypes * kMaxTypes];
  // per type: bit t set if an order-2 reaction pairs it with type t, and the
  // largest squared binding radius over those reactions (pair-search filters)
  uint32_t partnerMask[kMaxTypes];
  float maxBindR2[kMaxTypes];
  // geometry
  float lo[3], hi[3];
  float L[3], halfL[3];
  int32_t face[6];         // -x,+x,-y,+y,-z,+z
  int32_t periodic[3];
  float particleRadius;
  int32_t numSpheres;
  float sphere[kMaxSpheres][4];  // cx, cy, cz, R  (reflective from both sides)
  int32_t unbindOffsets;   // 1: order-1 splits place products at +-unbinding offsets
  // uniform (wrapped) hash grid (particles_kernel.cu:169-186)
  float origin[3];
  float cellSize[3];
  uint32_t gbits[3];       // log2 of the grid dimension per axis
  uint32_t bucketOff;      // first global bucket of this replica
  uint32_t numBuckets;     // 2^(gbits sum)
  // random streams
  uint32_t streamOff;      // first global stream (multiple of 64)
  uint32_t numStreams;     // S: particle i of this replica draws from stream i % S
  // order-0 reactions (smolreact.cpp:295-303, reaction_cpu.cpp:36-63): per
  // step, Poisson(prob) new particles of type prodX at uniform positions in
  // the box.  Poisson(prob) is drawn as the sum of zeroPieces Poisson
  // variates of mean prob / zeroPieces (each below 12), each by the
  // multiplication method of stats.cpp:poisrandF against zeroG = exp(-mean).
  int32_t numZero;
  int32_t zeroList[kMaxRxn];
  int32_t zeroPieces[kMaxRxn];
  float zeroG[kMaxRxn];
  uint32_t zeroMax;        // creations per step above this are an error
  uint32_t zeroOff;        // first slot of this replica in the creation buffer
};

}  // namespace smd
