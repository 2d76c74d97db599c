// Kernel argument blocks and launch wrappers (host-visible).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "smd_types.h"

namespace smd {

// Tile table of the segmented sort (k_seg_tiles): tiles of tileSize keys
// inside each replica's segment, laid out from the per-replica counts.
struct SegTilesArgs {
  const uint32_t* segOff;
  const uint32_t* segCount;
  int R;
  uint32_t tileSize;
  uint32_t maxTiles;
  uint2* tpos;
  uint32_t* tseg;
  uint32_t* nTiles;
  uint32_t* segFirst;
  uint32_t* nLong;
  uint32_t* repDead;
  const uint32_t* dN;
  uint32_t* status;
};

struct ProducerArgs {
  const ReplicaTables* tables;
  const int* streamReplica;   // per global stream: replica id (padding lanes of a replica included), -1 none
  const uint32_t* repOff;     // per replica: first particle
  const uint32_t* repCount;   // per replica: particle count
  uint32_t* mtState;          // [20][Stot]
  const uint32_t* mtPar;      // [3][Stot]
  uint32_t Stot;
  const float* gauss;         // entries 1 to 2048 of the 4096-entry table (gauss_entry)
  float4* pt;                 // canonical (x, y, z, type bits), updated in place; also the pair-search source
  float* w;                   // canonical fourth position component (set to 0 for offset products)
  float4* addedPos;
  int* birthType;
  uint32_t* key;
  uint32_t deadKey;
  uint32_t* repDeaths;
  uint32_t* repBirths;
  uint32_t* tileAlive;
  uint32_t* tileBirth;
  const uint32_t* retire;     // per replica: remove every particle this step (nullptr: none)
  // 1: one extra workgroup of the launch lays out the segmented sort's tiles
  // (it reads only the per-replica counts the step starts with)
  int segTilesOn = 0;
  SegTilesArgs segTiles;
};

struct SortBuffers {
  uint32_t* keys;
  uint32_t* vals;
  uint32_t* keysAlt;
  uint32_t* valsAlt;
  uint32_t* counts;  // sort_counts_words(maxN, radixBits)
  uint32_t* status;  // sticky status word (kStatusSortSpin), or nullptr
  int algo;          // 0: single sweep with look-back, 1: count / scan / scatter per digit
  // optional gather into sorted order: gatherDst[j] = gatherSrc[vals[j]] for
  // keys[j] < gatherDeadKey (fused into the last digit pass when possible)
  const float4* gatherSrc;
  float4* gatherDst;
  uint32_t gatherDeadKey;
  uint32_t sizeHint;  // expected key count (tile size choice only), 0: maxN
  // with the gather fused into the last pass (algo 1): 1 puts each particle's
  // index into its sorted record (w = index << kIndexShift | type, indices
  // below kMaxPackedIndex) instead of writing the sorted values
  int packIndex;
};

// Segments of an ensemble's replica-major keys for the segmented sort:
// segment r is [segOff[r], segOff[r] + segCount[r]), its keys are
// bucketOff[r] + (a local cell below 2^(2 bits - 1)) or deadKey.
struct SegSortArgs {
  int R;
  const uint32_t* segOff;
  const uint32_t* segCount;
  const uint32_t* bucketOff;
  uint32_t deadKey;
  int lowBits;         // first-pass digit width (7 to 10 bits)
  int highBits;        // second-pass digit width (7 to 10 bits)
  uint32_t* scratch;   // seg_sort_words(maxN, R, max(lowBits, highBits)) words
  int tilesLaid = 0;   // 1: the tile table is already laid out (seg_tiles_args, by the producer launch)
};

// Digit widths of the segmented sort for replica grids of at most 2^maxGridBits
// buckets: {low, high} with low + high > maxGridBits, or {0, 0} when two
// passes of at most 9 bits cannot cover them (the engine then sorts global
// keys).
void seg_sort_digits(uint32_t maxGridBits, int& lowBits, int& highBits);

struct CandArgs {
  const ReplicaTables* tables;
  const uint32_t* bucketStart;  // per replica bucketOff, [R] = deadKey
  const int* bucketRep;         // replica of bucket chunk key >> repShift (nullptr: search bucketStart)
  uint32_t repShift;
  int R;
  const uint32_t* keys;         // sorted
  const float4* sortedPT;
  const uint32_t* first;
  const uint32_t* dLive;
  int* cand;
};

struct ResolveArgs {
  const ReplicaTables* tables;
  const int* streamReplica;
  uint32_t* mtState;
  const uint32_t* mtPar;
  uint32_t Stot;
  const uint32_t* first;
  const int* cand;
  const float4* sortedPT;
  const uint32_t* vals;
  int packed;  // sortedPT carries the particles' indices (vals unused)
  float4* pt;
  float* w;
  uint32_t* repDeaths;
  uint32_t* tileAlive;
};

struct FinalizeArgs {
  int R;
  const uint32_t* repCount;
  uint32_t* repDeaths;
  uint32_t* repBirths;
  uint32_t* newOff;     // [R+1]
  uint32_t* newCount;
  uint32_t* aliveBase;
  uint32_t* birthBase;
  uint32_t* repAlive;
  uint32_t* dNnext;
  const uint32_t* dN;
  uint32_t* tileAlive;
  uint32_t* tileBirth;
  uint32_t* tileAlivePre;
  uint32_t* tileBirthPre;
  unsigned long long* particleSteps;  // += particles entering the step
  uint32_t* stepCount;                // steps completed (incremented here)
  unsigned long long* countRing;      // host-visible: slot (steps & ringMask) = steps << 32 | particles
  uint32_t ringMask;
  const uint32_t* zCount;  // per replica: particles created this step (nullptr: none)
  uint32_t* zeroBase;      // per replica: alive + births, where created particles go
  uint32_t* injCount;      // per replica: particles admitted this step (nullptr: none); reset
  uint32_t* injBase;       // per replica: alive + births + created, where admitted particles go
  uint32_t* retire;        // per replica retire flags (nullptr: none); reset
  uint32_t* blockSum;      // [2][blockStride]: alive and birth totals of each 1024-tile block
  uint32_t blockStride;    // at least finalize_blocks(maxN)
  uint32_t maxTiles;       // compaction tiles of the launch bound (set by launch_finalize)
};

struct ZeroArgs {
  const ReplicaTables* tables;
  int R;
  uint32_t* zMt;         // [20][R] one creation stream per replica
  const uint32_t* zPar;  // [3][R]
  float4* zPos;          // creation buffer (replica r at tables[r].zeroOff)
  int* zType;
  uint32_t* zCount;      // per replica: particles created this step
  uint32_t* status;      // sticky error bits (kStatusZeroOverflow)
};

struct InjectArgs {
  const float4* injPos;
  const int* injType;
  const int* injRep;          // replica of each admitted particle
  const uint32_t* injSrcOff;  // first admitted particle of each replica
  uint32_t nInj;
  const uint32_t* newOff;
  const uint32_t* injBase;
  float4* ptOut;
  float* wOut;
};

struct ScatterArgs {
  int R;
  const uint32_t* dN;
  const uint32_t* repOff;
  const uint32_t* newOff;
  const uint32_t* aliveBase;
  const uint32_t* birthBase;
  const uint32_t* repAlive;
  const uint32_t* tileAlivePre;
  const uint32_t* tileBirthPre;
  const float4* pt;
  const float* w;
  const float4* addedPos;
  int* birthType;          // nonzero only at births; cleared as they are taken
  uint32_t* tileBirth;     // births per compaction tile (producer); cleared here
  float4* ptOut;
  float* wOut;
  const uint32_t* blockSum;  // the finalize blocks' totals (FinalizeArgs::blockSum)
  uint32_t blockStride;
};

hipError_t launch_mt_init(uint32_t* st, const uint32_t* seeds, uint32_t S, hipStream_t stream);
hipError_t launch_mt_fill(uint32_t* st, const uint32_t* par, uint32_t S, uint32_t nStreams, int count, int batch3,
                          uint32_t* out, hipStream_t stream);
// blockSize: 128, 256 or 512 threads (identical results)
hipError_t launch_producer(const ProducerArgs& a, int blockSize, hipStream_t stream);
int sort_passes(uint32_t keyBits, int radixBits);
size_t sort_counts_words(uint32_t maxN, int radixBits);
hipError_t launch_sort(SortBuffers& b, const uint32_t* dN, uint32_t maxN, uint32_t keyBits, int radixBits,
                       bool identityVals, hipStream_t stream);
// Segmented two-pass sort of an ensemble's keys (same output as launch_sort
// with identityVals over all keys: live keys in order, removed ones after
// them in index order).
size_t seg_sort_words(uint32_t maxN, int R, int maxBits);
hipError_t launch_sort_seg(SortBuffers& b, const SegSortArgs& g, const uint32_t* dN, uint32_t maxN,
                           bool identityVals, hipStream_t stream);
// The tile-table arguments of a segmented sort of at most maxN keys (the
// addresses inside g.scratch that launch_sort_seg uses).
SegTilesArgs seg_tiles_args(const SegSortArgs& g, const uint32_t* dN, uint32_t maxN, uint32_t* status);
// Counter words the count / scan / scatter sort (SortBuffers::algo = 1)
// uses for maxN keys: per-tile digit counts and the digit totals.
size_t sort3k_counts_words(uint32_t maxN, uint32_t sizeHint, int radixBits);
// first: first_words(nb) words (first[0..nb]).  slotBits > 0 (slot engines,
// slots of 2^slotBits buckets): entries of slot s past offset
// tables[s].numBuckets are not written.
size_t first_words(uint32_t nb);
hipError_t launch_first(const uint32_t* keys, const uint32_t* dN, uint32_t maxN, uint32_t nb, uint32_t* first,
                        uint32_t* dLive, const ReplicaTables* tables, uint32_t slotBits, uint32_t nSlots,
                        hipStream_t stream);
hipError_t launch_gather(const uint32_t* keys, const uint32_t* vals, const float4* src, const uint32_t* dN,
                         uint32_t maxN, uint32_t deadKey, float4* sortedPT, hipStream_t stream);
// blockSize: 64, 128 or 256 threads (identical results)
hipError_t launch_candidates(const CandArgs& a, uint32_t maxN, int blockSize, hipStream_t stream);
hipError_t launch_resolve(const ResolveArgs& a, hipStream_t stream);
// Finalize workgroups for a step of at most maxN particles (one per 4096
// compaction tiles).
uint32_t finalize_blocks(uint32_t maxN);
hipError_t launch_finalize(const FinalizeArgs& a, uint32_t maxN, hipStream_t stream);
hipError_t launch_scatter(const ScatterArgs& a, uint32_t maxN, hipStream_t stream);
hipError_t launch_inject(const InjectArgs& a, hipStream_t stream);
hipError_t launch_zero(const ZeroArgs& a, hipStream_t stream);
// Appends replica r's created particles after its survivors and births.
hipError_t launch_zero_place(int R, const ReplicaTables* tables, const float4* zPos, const int* zType,
                             const uint32_t* zCount, const uint32_t* newOff, const uint32_t* zeroBase, float4* ptOut,
                             float* wOut, hipStream_t stream);
// pos4[i] = (pt[i].xyz, w[i]), type[i] = type bits of pt[i], i < n.
hipError_t launch_unpack(const float4* pt, const float* w, uint32_t n, float4* pos4, int* type, hipStream_t stream);
hipError_t launch_mt_init_range(uint32_t* st, const uint32_t* seeds, uint32_t Stot, uint32_t off, uint32_t count,
                                hipStream_t stream);

}  // namespace smd
