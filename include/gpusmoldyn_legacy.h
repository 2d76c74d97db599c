/*
 * The reference's own extern "C" layer, re-implemented for HIP on gfx950.
 *
 * Same function set, argument order and argument meaning as
 *   particleSystem.cuh:29-110   (and cudaGLInit, particleSystemCuda.cu:56-64)
 *   MersenneTwister.h:49        (InitializeRandomNumbers)
 *   radixsort.h:45-57           (radixSort)
 * so that ParticleSystem code written against the reference links against
 * libgpusmoldyn.so.  Differences, all deliberate:
 *   - every function returns an int status (0 = success, see smd_status in
 *     gpusmoldyn.h) instead of aborting the process; compactPosAndType keeps
 *     returning the compacted count (negative = -status);
 *   - OpenGL interop entry points exist but return SMD_ERR_UNSUPPORTED;
 *   - InitializeRandomNumbers seeds the 32768 streams deterministically
 *     (seed k = output k of the dcmt stream of record 0 seeded with 4357)
 *     instead of rand() after srand(time(0)) (MersenneTwister.cpp:133-135);
 *   - kernels that the reference runs with racing threads follow fixed
 *     rules: particle i draws from stream i % 32768 in increasing i; a
 *     particle reacts bimolecularly at most once (mutual nearest partner);
 *     compaction is order preserving and race free;
 *   - the CUDPP exclusive scan the reference calls around compactPosAndType
 *     (particleSystem.cpp:884, 888, 954) is smdExclusiveScanInt.
 * Structures keep the reference layouts (SimParams: particles_kernel.cuh:46-67,
 * MersenneTwisterState: MersenneTwister.h:43-47, Reaction: reaction.h:36-48).
 */
#ifndef GPUSMOLDYN_LEGACY_H_INCLUDED
#define GPUSMOLDYN_LEGACY_H_INCLUDED

#include <stddef.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

/* float3 / uint3 / int2 of the reference's signatures.  In C++ the first two
 * convert to any vector type with x, y, z members (HIP's float3 and uint3
 * among them), so code written against the reference's getters compiles
 * unchanged. */
typedef struct smd_f3 {
  float x, y, z;
#ifdef __cplusplus
  template <typename V, typename = decltype(((V*)0)->x + ((V*)0)->y + ((V*)0)->z)>
  operator V() const {
    V v;
    v.x = x;
    v.y = y;
    v.z = z;
    return v;
  }
#endif
} smd_f3;
typedef struct smd_u3 {
  unsigned int x, y, z;
#ifdef __cplusplus
  template <typename V, typename = decltype(((V*)0)->x + ((V*)0)->y + ((V*)0)->z)>
  operator V() const {
    V v;
    v.x = x;
    v.y = y;
    v.z = z;
    return v;
  }
#endif
} smd_u3;
typedef struct smd_i2 { int x, y; } smd_i2;

#ifdef __cplusplus
extern "C" {
#endif

/* particles_kernel.cuh:46-67 */
typedef struct SimParams {
  smd_f3 colliderPos;
  float colliderRadius;
  smd_f3 gravity;
  float globalDamping;
  float particleRadius;
  smd_u3 gridSize;
  unsigned int numCells;
  smd_f3 worldOrigin;
  smd_f3 cellSize;
  unsigned int numBodies;
  unsigned int maxParticlesPerCell;
  float spring;
  float damping;
  float shear;
  float attraction;
  float boundaryDamping;
} SimParams;

/* MersenneTwister.h:43-47 */
#define MT_RNG_COUNT 32768
typedef struct MersenneTwisterState {
  unsigned int mt[19];
  int iState;
  unsigned int mti1;
} MersenneTwisterState;

/* reaction.h:36-48: a device copy of this struct holds device pointers */
typedef struct Reaction {
  int* order;
  float* bindRadiusSquared;
  float* unbindRadius;
  float* geminateProbability;
  float* prob;
  float* rate;
  int* nprod;
  smd_i2* products;
  smd_i2* reactants;
  float* productPositions;
} Reaction;

/* reaction.h:52-66: the per-order reaction superstructure that
 * fillReactionOrderInfo (smolreact.cpp:173-271) builds (host arrays). */
typedef struct ReactionOrderInfo {
  int order;
  int reactionsPerOrder;
  int* reactionsPerTypeIdx;       /* where reactant type t starts in localToGlobal */
  int* localToGlobal;             /* reactions listed per reactant type (global ids) */
  int localToGlobalSize;
  int* numberOfReactionsPerType;  /* reactions per reactant type */
  int* reactionIndexes;           /* global ids of this order's reactions */
} ReactionOrderInfo;

int cudaInit(int argc, char** argv);
int cudaGLInit(int argc, char** argv);
int allocateArray(void** devPtr, size_t size);
int freeArray(void* devPtr);
int threadSync(void);
int copyArrayFromDevice(void* host, const void* device, unsigned int vbo, int size);
int copyArrayToDevice(void* device, const void* host, int offset, int size);
int registerGLBufferObject(unsigned int vbo);
int unregisterGLBufferObject(unsigned int vbo);
void* mapGLBufferObject(unsigned int vbo);
int unmapGLBufferObject(unsigned int vbo);

int setParameters(SimParams* hostParams);
int copyGaussianToConstant(float* h_gaussianLookupTable, unsigned int gaussianTableDim);
int initializeMT(MersenneTwisterState* stateArray);
void InitializeRandomNumbers(void);
/* Same, with an explicit master seed (InitializeRandomNumbers uses 4357). */
int smdInitializeRandomNumbersSeeded(unsigned int seed);

int integrateSystem(Reaction* reactionList, int* localToGlobalReaction1, int* reactionsPerTypeIdx1,
                    int* numberOfReactionsPerType1, int numberOfReactions1, float* pos, int* types, float* addedPos,
                    int* addedTypes, int* birthArray, int* deathArray, float* diffusionRates,
                    MersenneTwisterState* rngStateArray, float deltaTime, unsigned int numParticles,
                    unsigned int numTypes);
int adjustPositions(float* pos, int* types, unsigned int numParticles, float particleRadius);
int calcHash(unsigned int* gridParticleHash, unsigned int* gridParticleIndex, float* pos, int numParticles);
int reorderDataAndFindCellStart(unsigned int* cellStart, unsigned int* cellEnd, float* sortedPos,
                                unsigned int* gridParticleHash, unsigned int* gridParticleIndex, float* oldPos,
                                unsigned int numParticles, unsigned int numCells);
int collide(float* newPos, float* sortedPos, unsigned int* gridParticleIndex, int* types, int* birthDeath,
            float* diffusionCoefficients, int* reactionTable, Reaction* reactionList, unsigned int* cellStart,
            unsigned int* cellEnd, unsigned int numParticles, unsigned int numCells, unsigned int numTypes,
            MersenneTwisterState* rngStateArray);
int compactPosAndType(int* dCompactedTemp, int* deathBirth, float* dPos, int* dTypes, int numParticles);

/* radixsort.h:45-57: key/value sort in place; tempKeys/tempValues are n-element
 * device scratch; counters/countersSum/blockOffsets/scanPlan are unused
 * (the library keeps its own counters); flipBits must be false. */
int radixSort(unsigned int* keys, unsigned int* values, unsigned int* tempKeys, unsigned int* tempValues,
              unsigned int* counters, unsigned int* countersSum, unsigned int* blockOffsets, void* scanPlan,
              unsigned int numElements, unsigned int keyBits, bool flipBits);

/* No counterpart in the reference: radixSort for callers that cannot
 * allocate device memory (stream capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same sort, same arguments minus the unused ones; the
 * digit counters live in a fixed device array of the library (about 4 MiB per
 * GPU) instead of a buffer it allocates, so numElements is limited to 2^24
 * (larger: SMD_ERR_CAPACITY, nothing launched).  Calls that overlap on one
 * GPU share that array and race. */
int smdRadixSortNoAlloc(unsigned int* keys, unsigned int* values, unsigned int* tempKeys, unsigned int* tempValues,
                        unsigned int numElements, unsigned int keyBits);

/* Exclusive + scan of n ints on the device (out may alias in). */
int smdExclusiveScanInt(int* out, const int* in, unsigned int n);

#ifdef __cplusplus
}
#endif

#endif
