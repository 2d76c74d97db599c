// HIP kernels of the timestep for gfx950 (wave64).
//
// One step over an ensemble of R >= 1 replicas (a plain simulation is R = 1):
//   k_producer    diffuse + unimolecular reactions + boundaries + hash   (one thread per RNG stream)
//   k_sort_*      stable LSD radix sort of (bucket key, particle index); the last digit pass
//                 also copies each particle's (x, y, z, type) record into sorted order
//   k_first       lower bound of every bucket in the sorted keys
//   k_candidates  neighbour-cell search for each particle's nearest reactive partner
//   k_resolve     mutual-nearest pairs react (bimReact)                (one thread per RNG stream)
//   k_finalize    per-replica counts, offsets and compaction-tile prefixes (one workgroup)
//   k_scatter     order-preserving compaction of survivors, births appended per replica
// (k_gather copies the records after the single-sweep sort variant.)
#include <hip/hip_runtime.h>

#include <algorithm>

#include "smd_device.h"
#include "smd_kernels.h"

namespace smd {

constexpr int kBS = 256;  // workgroup size of the per-stream kernels

// ---------------------------------------------------------------------------
// RNG stream initialisation (MersenneTwister.cuh:40-49 / :80-91) with host
// supplied seeds.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mt_init(uint32_t* __restrict__ st, const uint32_t* __restrict__ seeds,
                                                 uint32_t S) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= S) return;
  uint32_t v = seeds[g];
  st[g] = v;
  for (int i = 1; i < kMtNN; ++i) {
    v = 1812433253u * (v ^ (v >> 30)) + (uint32_t)i;
    st[(size_t)i * S + g] = v;
  }
  st[(size_t)kMtNN * S + g] = 0u;
}

// Draws `count` numbers from each stream g < nStreams (out[g * count + k]),
// advancing the stored state: the stream-level parity hook.
__global__ __launch_bounds__(kBS) void k_mt_fill(uint32_t* __restrict__ st, const uint32_t* __restrict__ par,
                                                 uint32_t S, uint32_t nStreams, int count, int batch3,
                                                 uint32_t* __restrict__ out) {
  __shared__ uint32_t lds[kMtNN * kBS];
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nStreams) return;
  Mt m;
  mt_load<kBS>(m, &lds[threadIdx.x], st, par, S, g);
  int k = 0;
  if (batch3) {
    for (; k + 3 <= count; k += 3) {
      uint32_t a, b, c;
      mt_draw3<kBS>(m, a, b, c);
      out[(size_t)g * count + k] = a;
      out[(size_t)g * count + k + 1] = b;
      out[(size_t)g * count + k + 2] = c;
    }
  }
  for (; k < count; ++k) out[(size_t)g * count + k] = mt_draw<kBS>(m);
  mt_store<kBS>(m, st, S, g);
}

// ---------------------------------------------------------------------------
// Producer: diffuse (particles_kernel.cu:125-166) + uniReact
// (reaction_kernel.cuh:263-342) + boundaries (particles_kernel.cu:60-95 and
// the surface semantics of surfaces_kernel.cuh) + calcHash
// (particles_kernel.cu:188-207), fused.  Thread g owns stream g and handles
// the particles i = s, s + S, s + 2S, ... of its replica in increasing order,
// so every stream's draws happen in a fixed order.
// ---------------------------------------------------------------------------
// Per-type and per-reaction entries the producer indexes per particle, staged
// in LDS for each wave's replica (a few hundred bytes per wave).
struct ProducerLds {
  float rate[kMaxTypes];
  int32_t uniCount[kMaxTypes];
  int32_t uniStart[kMaxTypes];
  int32_t uniList[kMaxRxn];
  float prob[kMaxRxn];
  int32_t prodX[kMaxRxn];
  int32_t prodY[kMaxRxn];
};

// The segmented sort's tile table by one workgroup of BS threads (defined
// with k_seg_tiles below); wsum holds BS / 64 words, sLong one.
template <int BS>
__device__ __forceinline__ void seg_tiles_body(const SegTilesArgs& a, uint32_t* wsum, uint32_t* sLong);

// The replica tables arrive as a separate read-only, non-aliased argument:
// their wave-uniform fields then load through the scalar cache, whose waits
// do not hold up the particle records in flight.
template <int BS>
__global__ __launch_bounds__(BS) void k_producer(ProducerArgs a, const ReplicaTables* __restrict__ tables) {
  __shared__ uint32_t lds_mt[kMtNN * BS];
  __shared__ float lds_gauss[kGaussDim / 2];  // half of the table (gauss_entry)
  __shared__ ProducerLds lds_tab[BS / 64];
  if (a.segTilesOn && blockIdx.x == gridDim.x - 1u) {
    // the launch's extra workgroup: the segmented sort's tile table, which
    // needs only the per-replica counts this step starts with (no stream's
    // work depends on it, nor it on theirs)
    seg_tiles_body<BS>(a.segTiles, lds_mt, lds_mt + BS / 64);
    return;
  }
  const uint32_t g = blockIdx.x * BS + threadIdx.x;
  const uint32_t lane = threadIdx.x & 63u;
  const int r = (g < a.Stot) ? a.streamReplica[g] : -1;
  // a workgroup whose streams belong to no replica (the idle stream ranges of
  // a slot engine) leaves before it copies the Gaussian table
  if (!__syncthreads_or(r >= 0)) return;
  for (int k = threadIdx.x; k < kGaussDim / 2; k += BS) lds_gauss[k] = a.gauss[k];

  const int rw = __builtin_amdgcn_readfirstlane(r);  // a wave never spans two replicas
  ProducerLds& L = lds_tab[threadIdx.x >> 6];
  if (rw >= 0) {
    const ReplicaTables& T0 = tables[rw];
    if (lane < (uint32_t)kMaxTypes) {
      L.rate[lane] = T0.rate[lane];
      L.uniCount[lane] = T0.uniCount[lane];
      L.uniStart[lane] = T0.uniStart[lane];
    }
    if (lane < (uint32_t)kMaxRxn) {
      L.uniList[lane] = T0.uniList[lane];
      L.prob[lane] = T0.prob[lane];
      L.prodX[lane] = T0.prodX[lane];
      L.prodY[lane] = T0.prodY[lane];
    }
  }
  __syncthreads();
  uint32_t nDeaths = 0, nBirths = 0;

  if (rw >= 0) {
    const ReplicaTables& T = tables[rw];
    const uint32_t S = T.numStreams;
    const uint32_t s = g - T.streamOff;
    const uint32_t s0 = (g & ~63u) - T.streamOff;  // stream of lane 0
    const bool laneOK = (r >= 0) && (s < S);
    const uint32_t n = a.repCount[rw];
    const uint32_t off = a.repOff[rw];
    const uint32_t bucketOff = T.bucketOff;
    // A retiring replica (a finished serving job) has all its particles
    // removed this step, without drawing.
    const bool kill = a.retire != nullptr && a.retire[rw] != 0u;
    const uint32_t iters = (n > s0) ? (n - s0 + S - 1) / S : 0;
    // the replica's box, grid and reaction flags in registers for the loop
    const BoxGeom G = box_geom(T);
    const int hasUni = T.hasUni, unbindOffsets = T.unbindOffsets, numSpheres = T.numSpheres;
    // Records are requested one pair of particles ahead.
    auto rec = [&](uint32_t it) __attribute__((always_inline)) {
      const uint32_t iq = s + it * S;
      return (laneOK && iq < n) ? a.pt[off + iq] : make_float4(0.f, 0.f, 0.f, 0.f);
    };
    float4 p0 = rec(0), p1 = rec(1);
    Mt m;
    if (laneOK && !kill) mt_load<BS>(m, &lds_mt[threadIdx.x], a.mtState, a.mtPar, a.Stot, g);
    // the stream parameters are complete before the loop, so no wait inside
    // it covers them (such a wait would also cover the records in flight)
    asm volatile("" ::"v"(m.a), "v"(m.mb), "v"(m.mc), "v"(m.mti1));
    // returns the particle's births << 32 | deaths
    auto step_one = [&](const float4 pCur, const uint32_t it) __attribute__((always_inline)) -> uint64_t {
      uint64_t db = 0;
      const uint32_t il = s + it * S;
      const bool act = laneOK && il < n;
      const uint32_t i = off + il;
      bool alive = false, birth = false;
      if (act && kill) {
        a.pt[i] = make_float4(pCur.x, pCur.y, pCur.z, __int_as_float(0));
        a.key[i] = a.deadKey;
        db += 1u;
      } else if (act) {
        const float4 p = pCur;
        int t = __float_as_int(pCur.w);
        uint32_t u0, u1, u2;
        mt_draw3<BS>(m, u0, u1, u2);
        const float rt = L.rate[t > 0 ? t : 0];
        float np[3];
        np[0] = fmaf(rt, gauss_entry(lds_gauss, u0 & (kGaussDim - 1)), p.x);
        np[1] = fmaf(rt, gauss_entry(lds_gauss, u1 & (kGaussDim - 1)), p.y);
        np[2] = fmaf(rt, gauss_entry(lds_gauss, u2 & (kGaussDim - 1)), p.z);
        bool wZero = false;  // the fourth component becomes 0 (offset products)
        int bt = 0;
        float bp[3] = {np[0], np[1], np[2]};
        float bw = 0.0f;
        if (t <= 0) {
          t = 0;
        } else if (hasUni) {
          const int cnt = L.uniCount[t];
          const int base = L.uniStart[t];
          for (int j = 0; j < cnt; ++j) {
            const int rx = L.uniList[base + j];
            if (rand_real(mt_draw<BS>(m)) < L.prob[rx]) {
              if (L.prodX[rx] <= 0) {
                t = 0;
              } else {
                t = L.prodX[rx];
                const int t2 = L.prodY[rx];
                if (t2 > 0) {
                  bt = t2;
                  if (unbindOffsets) {
                    const float th = smd_theta(rand_real(mt_draw<BS>(m)));
                    const float ph = rand_real(mt_draw<BS>(m)) * 6.28318548f;
                    const float ch = rand_real(mt_draw<BS>(m)) * 6.28318548f;
                    float v1[3], v2[3];
                    rotate_offsets(th, ph, ch, T.prodPos[rx], v1, v2);
                    bp[0] = np[0] + v2[0];
                    bp[1] = np[1] + v2[1];
                    bp[2] = np[2] + v2[2];
                    np[0] = np[0] + v1[0];
                    np[1] = np[1] + v1[1];
                    np[2] = np[2] + v1[2];
                    wZero = true;
                  } else {
                    bw = a.w[i];  // the birth inherits the parent's fourth component
                  }
                }
              }
              break;
            }
          }
        }
        uint32_t key = a.deadKey;
        if (t > 0) {
          const float o[3] = {p.x, p.y, p.z};
          if (!apply_box(G, np)) {
            t = 0;
          } else {
            if (numSpheres) apply_spheres(T, o, np);
            key = bucketOff + grid_hash(G, np[0], np[1], np[2]);
          }
        }
        a.pt[i] = make_float4(np[0], np[1], np[2], __int_as_float(t));
        if (wZero) a.w[i] = 0.0f;
        a.key[i] = key;
        // birth types are zero except where the scatter has a birth to take
        // (it clears them again), so only births are written
        if (bt) {
          a.birthType[i] = bt;
          a.addedPos[i] = make_float4(bp[0], bp[1], bp[2], bw);
        }
        alive = t > 0;
        birth = bt > 0;
        db += (alive ? 0ull : 1ull) + (birth ? (1ull << 32) : 0ull);
      }
      // Compaction tile counters: lanes hold consecutive particle indices.
      const uint64_t am = __ballot(alive);
      const uint64_t bm = __ballot(birth);
      if (lane == 0) {
        const uint32_t base = off + s0 + it * S;
        const uint32_t t0 = base / kCompactTile;
        const uint32_t split = (t0 + 1) * kCompactTile - base;  // first lane in the next tile
        const uint64_t loMask = split >= 64 ? ~0ull : ((1ull << split) - 1ull);
        const uint32_t aLo = __popcll(am & loMask), aHi = __popcll(am & ~loMask);
        const uint32_t bLo = __popcll(bm & loMask), bHi = __popcll(bm & ~loMask);
        if (aLo) atomicAdd(&a.tileAlive[t0], aLo);
        if (aHi) atomicAdd(&a.tileAlive[t0 + 1], aHi);
        if (bLo) atomicAdd(&a.tileBirth[t0], bLo);
        if (bHi) atomicAdd(&a.tileBirth[t0 + 1], bHi);
      }
      return db;
    };
    uint32_t it = 0;
    uint64_t db = 0;
    for (; it + 2 <= iters; it += 2) {
      // this pair's records were requested one pair ago; the next pair's are
      // requested before this pair's stores, so waiting for them later does
      // not wait for those stores
      const float4 c0 = p0, c1 = p1;
      p0 = rec(it + 2);
      p1 = rec(it + 3);
      db += step_one(c0, it);
      db += step_one(c1, it + 1);
    }
    if (it < iters) db += step_one(p0, it);
    nDeaths += (uint32_t)db;
    nBirths += (uint32_t)(db >> 32);
    if (laneOK && !kill) mt_store<BS>(m, a.mtState, a.Stot, g);
  }
  nDeaths = wave_sum(nDeaths);
  nBirths = wave_sum(nBirths);
  if (lane == 0 && rw >= 0) {
    if (nDeaths) atomicAdd(&a.repDeaths[rw], nDeaths);
    if (nBirths) atomicAdd(&a.repBirths[rw], nBirths);
  }
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort of (key, value), RADIX_BITS per pass, 4096-element
// tiles: upsweep (tile digit counts) -> rowscan (per digit across tiles) ->
// downsweep (wave64 ballot match ranking, LDS staging, scatter).  The
// result equals std::stable_sort by key of (key, index) pairs, the contract
// of RadixSort::sort (radixsort.h:168-185, radixsort.cu:1045-1142).
// ---------------------------------------------------------------------------
constexpr int kSortItems = 16;
constexpr int kSortTile = 256 * kSortItems;

// Tile of workgroup b when workgroups are dealt round robin over the 8 XCDs:
// each XCD gets a contiguous run of tiles, so neighbouring tiles (whose
// output runs for one digit are adjacent) share an L2 and their partial
// lines merge there.  A bijection on [0, nTiles).
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t nTiles) {
  const uint32_t x = b & 7u, j = b >> 3, q = nTiles >> 3, r = nTiles & 7u;
  return x * q + min(x, r) + j;
}

template <int BITS, int ITEMS>
__global__ __launch_bounds__(256) void k_sort_up(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ dN,
                                                 uint32_t* __restrict__ counts, uint32_t stride, int shift) {
  constexpr int RADIX = 1 << BITS;
  constexpr uint32_t TILE = 256u * ITEMS;
  __shared__ uint32_t hist[RADIX];
  const uint32_t n = min(*dN, (stride - 1u) * TILE);  // never beyond the sized grid
  const uint32_t nTiles = (n + TILE - 1) / TILE;
  if (blockIdx.x >= nTiles) return;
  const uint32_t tile = xcd_tile(blockIdx.x, nTiles);
  const uint32_t base = tile * TILE;
  uint32_t kk[ITEMS];
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const uint32_t e = base + k * 256 + threadIdx.x;
    kk[k] = e < n ? keys[e] : 0xFFFFFFFFu;
  }
  for (int d = threadIdx.x; d < RADIX; d += 256) hist[d] = 0;
  __syncthreads();
  // A chunk of 64 keys whose digits agree (the high digits of cell-ordered
  // or replica-ordered keys mostly do) adds its count with one LDS atomic
  // instead of 64 on the same address.
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const bool valid = base + k * 256 + threadIdx.x < n;
    const uint32_t d = (kk[k] >> shift) & (RADIX - 1);
    const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
    const uint64_t vm = __ballot(valid);
    if (__all(d == d0 || !valid)) {
      if ((threadIdx.x & 63u) == 0u && vm) atomicAdd(&hist[d0], (uint32_t)__popcll(vm));
    } else if (valid) {
      atomicAdd(&hist[d], 1u);
    }
  }
  __syncthreads();
  for (int d = threadIdx.x; d < RADIX; d += 256) counts[(size_t)d * stride + tile] = hist[d];
}

// Exclusive scan of 256 values held one per thread; returns the exclusive
// prefix and writes the total to *total (valid after the call).
__device__ __forceinline__ uint32_t block_exscan256(uint32_t v, uint32_t* wsum, uint32_t* total) {
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  uint32_t pre = 0;
  for (uint32_t k = 0; k < w; ++k) pre += wsum[k];
  if (threadIdx.x == 255) *total = pre + x;
  __syncthreads();
  return pre + x - v;
}

template <int BITS, int ITEMS>
__global__ __launch_bounds__(256) void k_sort_rowscan(uint32_t* __restrict__ counts, uint32_t* __restrict__ totals,
                                                      const uint32_t* __restrict__ dN, uint32_t stride) {
  constexpr uint32_t TILE = 256u * ITEMS;
  constexpr int PER = 8;  // consecutive tile counters per thread
  __shared__ uint32_t wsum[4];
  __shared__ uint32_t tot;
  const uint32_t n = min(*dN, (stride - 1u) * TILE);
  const uint32_t nTiles = (n + TILE - 1) / TILE;
  const uint32_t d = blockIdx.x;
  uint32_t* row = counts + (size_t)d * stride;
  uint32_t carry = 0;
  for (uint32_t b = 0; b < nTiles; b += 256 * PER) {
    const uint32_t e0 = b + threadIdx.x * PER;
    uint32_t v[PER], sum = 0;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      v[i] = (e0 + i < nTiles) ? row[e0 + i] : 0u;
      sum += v[i];
    }
    uint32_t run = carry + block_exscan256(sum, wsum, &tot);
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      if (e0 + i < nTiles) row[e0 + i] = run;
      run += v[i];
    }
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) totals[d] = carry;
}

// The same row scan with one wave per digit (four digits per workgroup) and
// no barriers, for small sorts, where a row is a fraction of one round of
// 512 counters and the workgroup barriers were most of the time.
template <int BITS, int ITEMS>
__global__ __launch_bounds__(256) void k_sort_rowscan_wave(uint32_t* __restrict__ counts,
                                                           uint32_t* __restrict__ totals,
                                                           const uint32_t* __restrict__ dN, uint32_t stride) {
  constexpr uint32_t TILE = 256u * ITEMS;
  constexpr int PER = 8;
  const uint32_t n = min(*dN, (stride - 1u) * TILE);
  const uint32_t nTiles = (n + TILE - 1) / TILE;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t d = blockIdx.x * 4u + (threadIdx.x >> 6);
  if (d >= (1u << BITS)) return;
  uint32_t* row = counts + (size_t)d * stride;
  uint32_t carry = 0;
  for (uint32_t b = 0; b < nTiles; b += 64 * PER) {
    const uint32_t e0 = b + lane * PER;
    uint32_t v[PER], sum = 0;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      v[i] = (e0 + i < nTiles) ? row[e0 + i] : 0u;
      sum += v[i];
    }
    uint32_t x = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o, 64);
      if (lane >= (uint32_t)o) x += y;
    }
    uint32_t run = carry + x - sum;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      if (e0 + i < nTiles) row[e0 + i] = run;
      run += v[i];
    }
    carry += __shfl(x, 63, 64);
  }
  if (lane == 0) totals[d] = carry;
}

template <int BITS, int ITEMS, bool GATHER>
__global__ __launch_bounds__(256) void k_sort_down(const uint32_t* __restrict__ keysIn,
                                                   const uint32_t* __restrict__ valsIn,
                                                   uint32_t* __restrict__ keysOut, uint32_t* __restrict__ valsOut,
                                                   const uint32_t* __restrict__ counts,
                                                   const uint32_t* __restrict__ totals,
                                                   const uint32_t* __restrict__ dN, uint32_t stride, int shift,
                                                   int identityVals, const float4* __restrict__ gatherSrc,
                                                   float4* __restrict__ gatherDst, uint32_t gatherDeadKey,
                                                   int packIndex) {
  constexpr int RADIX = 1 << BITS;
  constexpr int DPT = (RADIX + 255) / 256;  // digits per thread in the per-digit loops
  constexpr uint32_t TILE = 256u * ITEMS;
  __shared__ uint32_t s_base[RADIX];
  __shared__ uint32_t s_wcnt[4][RADIX];
  __shared__ uint32_t s_tstart[RADIX];
  __shared__ uint32_t s_key[TILE];
  __shared__ uint32_t s_val[TILE];
  __shared__ uint32_t wsum[4];
  __shared__ uint32_t tot;

  const uint32_t n = min(*dN, (stride - 1u) * TILE);
  const uint32_t nTiles = (n + TILE - 1) / TILE;
  if (blockIdx.x >= nTiles) return;
  const uint32_t tile = xcd_tile(blockIdx.x, nTiles);
  const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
  // the tile's keys and values are requested first: they arrive while the
  // digit starts are computed
  const uint32_t base = tile * TILE + w * (64 * ITEMS);
  uint32_t kk[ITEMS], vv[ITEMS], rk[ITEMS];
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    const uint32_t e = base + it * 64 + lane;
    const bool valid = e < n;
    kk[it] = valid ? keysIn[e] : 0u;
    vv[it] = identityVals ? e : (valid ? valsIn[e] : 0u);
  }

  // Global start of each digit for this tile; thread t owns the DPT
  // consecutive digits from t * DPT (one block scan whatever the radix).
  {
    uint32_t v[DPT], sum = 0;
#pragma unroll
    for (int q = 0; q < DPT; ++q) {
      const uint32_t d = t * DPT + q;
      v[q] = (d < (uint32_t)RADIX) ? totals[d] : 0u;
      sum += v[q];
    }
    uint32_t run = block_exscan256(sum, wsum, &tot);
#pragma unroll
    for (int q = 0; q < DPT; ++q) {
      const uint32_t d = t * DPT + q;
      if (d < (uint32_t)RADIX) s_base[d] = run + counts[(size_t)d * stride + tile];
      run += v[q];
    }
    __syncthreads();
  }
  for (uint32_t d = t; d < (uint32_t)RADIX; d += 256) {
    s_wcnt[0][d] = 0;
    s_wcnt[1][d] = 0;
    s_wcnt[2][d] = 0;
    s_wcnt[3][d] = 0;
  }
  __syncthreads();

#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    const uint32_t e = base + it * 64 + lane;
    const bool valid = e < n;
    const uint32_t k = kk[it];
    const uint32_t d = (k >> shift) & (RADIX - 1);
    const uint64_t mm = wave_digit_peers<BITS>(d, valid);
    // every lane of the digit writes the same new count (a lane past the
    // end counts its valid peers only, which is the same number)
    const uint32_t prev = s_wcnt[w][d];
    rk[it] = prev + lanes_below(mm);
    s_wcnt[w][d] = prev + (uint32_t)__popcll(mm);
  }
  __syncthreads();
  // Exclusive scan over waves per digit, and over digits for the tile (the
  // same digit ownership, one block scan).
  {
    uint32_t c[DPT][4], tsum = 0;
#pragma unroll
    for (int q = 0; q < DPT; ++q) {
      const uint32_t d = t * DPT + q;
#pragma unroll
      for (int k = 0; k < 4; ++k) c[q][k] = (d < (uint32_t)RADIX) ? s_wcnt[k][d] : 0u;
      tsum += c[q][0] + c[q][1] + c[q][2] + c[q][3];
    }
    uint32_t run = block_exscan256(tsum, wsum, &tot);
#pragma unroll
    for (int q = 0; q < DPT; ++q) {
      const uint32_t d = t * DPT + q;
      if (d < (uint32_t)RADIX) {
        s_wcnt[0][d] = 0;
        s_wcnt[1][d] = c[q][0];
        s_wcnt[2][d] = c[q][0] + c[q][1];
        s_wcnt[3][d] = c[q][0] + c[q][1] + c[q][2];
        s_tstart[d] = run;
      }
      run += c[q][0] + c[q][1] + c[q][2] + c[q][3];
    }
    __syncthreads();
  }
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    const uint32_t e = base + it * 64 + lane;
    if (e < n) {
      const uint32_t d = (kk[it] >> shift) & (RADIX - 1);
      const uint32_t lp = s_tstart[d] + s_wcnt[w][d] + rk[it];
      s_key[lp] = kk[it];
      s_val[lp] = vv[it];
    }
  }
  __syncthreads();
  const uint32_t tileN = min(TILE, n - tile * TILE);
  if (GATHER) {
    // every record load of the thread in flight before the first store
    uint32_t dk[ITEMS], dv[ITEMS], dd[ITEMS];
    float4 rec[ITEMS];
#pragma unroll
    for (int q = 0; q < ITEMS; ++q) {
      const uint32_t idx = t + q * 256u;
      const bool ok = idx < tileN;
      const uint32_t k = ok ? s_key[idx] : 0xFFFFFFFFu;
      const uint32_t d = (k >> shift) & (RADIX - 1);
      dk[q] = k;
      dv[q] = ok ? s_val[idx] : 0u;
      dd[q] = ok ? s_base[d] + (idx - s_tstart[d]) : 0u;
      rec[q] = (ok && k < gatherDeadKey) ? gatherSrc[dv[q]] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int q = 0; q < ITEMS; ++q) {
      if (t + q * 256u < tileN) {
        keysOut[dd[q]] = dk[q];
        if (packIndex) {
          // the record carries its particle's index above the type bits
          if (dk[q] < gatherDeadKey)
            gatherDst[dd[q]] = make_float4(rec[q].x, rec[q].y, rec[q].z,
                                           __int_as_float(pack_index_type(dv[q], __float_as_int(rec[q].w))));
        } else {
          valsOut[dd[q]] = dv[q];
          if (dk[q] < gatherDeadKey) gatherDst[dd[q]] = rec[q];
        }
      }
    }
    return;
  }
  for (uint32_t idx = t; idx < tileN; idx += 256) {
    const uint32_t k = s_key[idx];
    const uint32_t d = (k >> shift) & (RADIX - 1);
    const uint32_t dst = s_base[d] + (idx - s_tstart[d]);
    keysOut[dst] = k;
    valsOut[dst] = s_val[idx];
  }
}

// ---------------------------------------------------------------------------
// Single-sweep variant: one histogram kernel for all digit passes, then one
// binning kernel per pass.  Tiles are numbered in the order workgroups start
// (atomic counter); each tile publishes its per-digit count and finds the
// exclusive prefix over earlier tiles by decoupled look-back on per-tile
// status words (bit 31: inclusive prefix, bit 30: tile count only, bits
// 0-29: value).  A tile waits only on tiles numbered before it, which are
// already running, and every wait is bounded (error flag, no hang).
// ---------------------------------------------------------------------------
constexpr uint32_t kOsPrefix = 0x80000000u;
constexpr uint32_t kOsAgg = 0x40000000u;
constexpr uint32_t kOsValue = 0x3FFFFFFFu;
constexpr uint32_t kOsSpinLimit = 1u << 22;

__global__ __launch_bounds__(256) void k_os_zero(uint32_t* p, uint32_t n) {
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) p[i] = 0u;
}

template <int BITS>
__global__ __launch_bounds__(256) void k_os_hist(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ dN,
                                                 uint32_t maxN, int passes, uint32_t* __restrict__ hist) {
  constexpr int RADIX = 1 << BITS;
  __shared__ uint32_t h[(32 / BITS) * RADIX];
  const uint32_t n = min(*dN, maxN);
  for (int k = threadIdx.x; k < passes * RADIX; k += 256) h[k] = 0u;
  __syncthreads();
  for (uint32_t e = blockIdx.x * 256 + threadIdx.x; e < n; e += gridDim.x * 256) {
    const uint32_t k = keys[e];
    for (int p = 0; p < passes; ++p) atomicAdd(&h[p * RADIX + ((k >> (p * BITS)) & (RADIX - 1))], 1u);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < passes * RADIX; k += 256)
    if (h[k]) atomicAdd(&hist[k], h[k]);
}

template <int BITS>
__global__ __launch_bounds__(256) void k_os_pass(const uint32_t* __restrict__ keysIn,
                                                 const uint32_t* __restrict__ valsIn, uint32_t* __restrict__ keysOut,
                                                 uint32_t* __restrict__ valsOut, const uint32_t* __restrict__ dN,
                                                 uint32_t maxN, const uint32_t* __restrict__ hist,
                                                 uint32_t* status, uint32_t* tileCounter, uint32_t* err, int shift,
                                                 int identityVals) {
  constexpr int RADIX = 1 << BITS;
  constexpr int DPT = (RADIX + 255) / 256;
  __shared__ uint32_t s_base[RADIX];
  __shared__ uint32_t s_wcnt[4][RADIX];
  __shared__ uint32_t s_tstart[RADIX];
  __shared__ uint32_t s_key[kSortTile];
  __shared__ uint32_t s_val[kSortTile];
  __shared__ uint32_t wsum[4];
  __shared__ uint32_t tot;
  __shared__ uint32_t s_tile;

  const uint32_t n = min(*dN, maxN);
  const uint32_t nTiles = (n + kSortTile - 1) / kSortTile;
  if (threadIdx.x == 0) s_tile = atomicAdd(tileCounter, 1u);
  const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
  for (uint32_t d = t; d < (uint32_t)RADIX; d += 256) {
    s_wcnt[0][d] = 0;
    s_wcnt[1][d] = 0;
    s_wcnt[2][d] = 0;
    s_wcnt[3][d] = 0;
  }
  __syncthreads();
  const uint32_t tile = s_tile;
  if (tile >= nTiles) return;

  // local ranks (wave64 ballot match per 64-element chunk, element order)
  const uint32_t base = tile * kSortTile + w * (64 * kSortItems);
  uint32_t kk[kSortItems], vv[kSortItems], rk[kSortItems];
#pragma unroll
  for (int it = 0; it < kSortItems; ++it) {
    const uint32_t e = base + it * 64 + lane;
    const bool valid = e < n;
    const uint32_t k = valid ? keysIn[e] : 0u;
    kk[it] = k;
    vv[it] = identityVals ? e : (valid ? valsIn[e] : 0u);
    const uint32_t d = (k >> shift) & (RADIX - 1);
    const uint64_t mm = wave_digit_peers<BITS>(d, valid);
    // every lane of the digit writes the same new count (a lane past the
    // end counts its valid peers only, which is the same number)
    const uint32_t prev = s_wcnt[w][d];
    rk[it] = prev + lanes_below(mm);
    s_wcnt[w][d] = prev + (uint32_t)__popcll(mm);
  }
  __syncthreads();
  // per-digit wave offsets, tile counts, look-back, global bases
  {
    uint32_t carry = 0, gcarry = 0;
#pragma unroll
    for (int q = 0; q < DPT; ++q) {
      const uint32_t d = q * 256 + t;
      const bool dv = d < (uint32_t)RADIX;
      uint32_t c0 = 0, c1 = 0, c2 = 0, c3 = 0, g = 0;
      if (dv) {
        c0 = s_wcnt[0][d];
        c1 = s_wcnt[1][d];
        c2 = s_wcnt[2][d];
        c3 = s_wcnt[3][d];
        g = hist[d];
      }
      const uint32_t cnt = c0 + c1 + c2 + c3;
      uint32_t excl = 0;
      if (dv) {
        uint32_t* st = status + (size_t)tile * RADIX + d;
        if (tile == 0) {
          __hip_atomic_store(st, kOsPrefix | cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
          __hip_atomic_store(st, kOsAgg | cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          int32_t tp = (int32_t)tile - 1;
          uint32_t spins = 0;
          while (tp >= 0) {
            const uint32_t v =
                __hip_atomic_load(status + (size_t)tp * RADIX + d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((v & (kOsPrefix | kOsAgg)) == 0u) {
              if (++spins > kOsSpinLimit) {
                atomicOr(err, kStatusSortSpin);
                break;
              }
              __builtin_amdgcn_s_sleep(1);
              continue;
            }
            excl += v & kOsValue;
            if (v & kOsPrefix) break;
            --tp;
          }
          __hip_atomic_store(st, kOsPrefix | (excl + cnt), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      const uint32_t ex = block_exscan256(cnt, wsum, &tot);
      const uint32_t tt = tot;
      __syncthreads();
      const uint32_t gex = block_exscan256(g, wsum, &tot);
      const uint32_t gt = tot;
      if (dv) {
        s_wcnt[0][d] = 0;
        s_wcnt[1][d] = c0;
        s_wcnt[2][d] = c0 + c1;
        s_wcnt[3][d] = c0 + c1 + c2;
        s_tstart[d] = carry + ex;
        s_base[d] = gcarry + gex + excl;
      }
      carry += tt;
      gcarry += gt;
      __syncthreads();
    }
  }
#pragma unroll
  for (int it = 0; it < kSortItems; ++it) {
    const uint32_t e = base + it * 64 + lane;
    if (e < n) {
      const uint32_t d = (kk[it] >> shift) & (RADIX - 1);
      const uint32_t lp = s_tstart[d] + s_wcnt[w][d] + rk[it];
      s_key[lp] = kk[it];
      s_val[lp] = vv[it];
    }
  }
  __syncthreads();
  const uint32_t tileN = min((uint32_t)kSortTile, n - tile * kSortTile);
  for (uint32_t idx = t; idx < tileN; idx += 256) {
    const uint32_t k = s_key[idx];
    const uint32_t d = (k >> shift) & (RADIX - 1);
    const uint32_t dst = s_base[d] + (idx - s_tstart[d]);
    keysOut[dst] = k;
    valsOut[dst] = s_val[idx];
  }
}

// ---------------------------------------------------------------------------
// first[h] = lower bound of bucket h in the sorted keys, for h in [0, nb)
// (nb = deadKey + 1, so first[deadKey] = number of live particles).  This is
// the start/end table of reorderDataAndFindCellStartD
// (particles_kernel.cu:209-272) written for every bucket, so no memset is
// needed and bucket run [a, b] is [first[a], first[b + 1]).
// ---------------------------------------------------------------------------
// A workgroup covers kFirstSpan consecutive buckets [h0, h1).  Two waves find
// the span's key range [lo, hi) (lower bounds of h0 and h1) by 64-way
// searches; every particle j of the range that starts a run of its key marks
// bucket keys[j] with j, and a suffix minimum over the span (unmarked buckets
// take the next mark, or hi) gives every bucket's lower bound.  Reads the
// sorted keys once and writes first[] once, both coalesced, with the same
// work per workgroup whatever the key distribution.
constexpr uint32_t kFirstSpan = 4096;
constexpr uint32_t kFirstPer = kFirstSpan / 256;

// LDS index with one pad word per 32: a thread's 16 consecutive entries and
// its neighbours' fall in different banks.
__device__ __forceinline__ uint32_t first_pad(uint32_t i) { return i + (i >> 5); }

// Lower bound of v in the sorted keys[0, n), by one whole wave: each round
// loads 64 evenly spaced pivots and keeps the gap the answer lies in.
__device__ __forceinline__ uint32_t wave_lower_bound(const uint32_t* __restrict__ keys, uint32_t n, uint32_t v,
                                                     uint32_t lane) {
  uint32_t lo = 0, hi = n;  // the answer a satisfies lo <= a <= hi
  while (hi - lo > 64u) {
    const uint32_t step = (hi - lo + 63u) / 64u;
    const uint32_t x = lo + (lane + 1u) * step - 1u;
    const bool less = x < hi && keys[x] < v;  // true on a prefix of the lanes
    const uint32_t c = (uint32_t)__popcll(__ballot(less));
    const uint32_t nhi = min(lo + (c + 1u) * step - 1u, hi);
    lo += c * step;
    hi = nhi;
  }
  const uint32_t x = lo + lane;
  const bool less = x < hi && keys[x] < v;
  return lo + (uint32_t)__popcll(__ballot(less));
}

// Slot engines give every slot a fixed range of 2^slotBits buckets, of which
// its job's grid uses the first tables[slot].numBuckets (0 for a slot never
// used).  No key falls past that and the pair search and resolve read the
// table up to offset numBuckets, so spans that start beyond it are skipped
// (slotBits = 0: every span is written).
__global__ __launch_bounds__(256) void k_first(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ dN,
                                               uint32_t maxN, uint32_t nb, uint32_t* __restrict__ first,
                                               uint32_t* __restrict__ dLive, const ReplicaTables* __restrict__ tables,
                                               uint32_t slotBits, uint32_t nSlots) {
  __shared__ uint32_t s_out[kFirstSpan + kFirstSpan / 32];
  __shared__ uint32_t s_rng[2];
  __shared__ uint32_t s_wmin[4];
  const uint32_t h0 = blockIdx.x * kFirstSpan;
  if (slotBits) {
    const uint32_t slot = h0 >> slotBits;
    if (slot < nSlots && (h0 & ((1u << slotBits) - 1u)) > tables[slot].numBuckets) return;
  }
  const uint32_t n = min(*dN, maxN);
  const uint32_t h1 = min(h0 + kFirstSpan, nb);
  const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
  if (w == 0) {
    const uint32_t r = wave_lower_bound(keys, n, h0, lane);
    if (lane == 0) s_rng[0] = r;
  } else if (w == 1) {
    const uint32_t r = h1 == nb ? n : wave_lower_bound(keys, n, h1, lane);
    if (lane == 0) s_rng[1] = r;
  }
  for (uint32_t i = t; i < kFirstSpan; i += 256) s_out[first_pad(i)] = 0xFFFFFFFFu;
  __syncthreads();
  const uint32_t lo = s_rng[0], hi = max(s_rng[1], lo);
  // eight keys (and their predecessors) per thread in flight per round
  constexpr int kMarkPer = 8;
  for (uint32_t j0 = lo + t; j0 < hi; j0 += 256u * kMarkPer) {
    uint32_t kk[kMarkPer], kp[kMarkPer];
#pragma unroll
    for (int q = 0; q < kMarkPer; ++q) {
      const uint32_t j = j0 + (uint32_t)q * 256u;
      kk[q] = j < hi ? keys[j] : 0u;
      kp[q] = (j < hi && j > lo) ? keys[j - 1] : 0xFFFFFFFFu;
    }
#pragma unroll
    for (int q = 0; q < kMarkPer; ++q) {
      const uint32_t j = j0 + (uint32_t)q * 256u;
      const uint32_t k = kk[q];
      if (j < hi && k >= h0 && k < h1 && (j == lo || kp[q] != k)) s_out[first_pad(k - h0)] = j;
    }
  }
  __syncthreads();
  // suffix minimum: own 16 entries, then the minimum over later threads
  uint32_t v[kFirstPer];
  uint32_t m = hi;
#pragma unroll
  for (int i = kFirstPer - 1; i >= 0; --i) {
    m = min(m, s_out[first_pad(t * kFirstPer + i)]);
    v[i] = m;
  }
  uint32_t x = m;  // becomes the minimum over lanes >= lane
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_down(x, o, 64);
    if (lane + (uint32_t)o < 64u) x = min(x, y);
  }
  if (lane == 0) s_wmin[w] = x;
  const uint32_t xNext = __shfl_down(x, 1, 64);
  __syncthreads();
  uint32_t right = lane == 63u ? hi : xNext;
  for (uint32_t q = w + 1; q < 4u; ++q) right = min(right, s_wmin[q]);
#pragma unroll
  for (int i = 0; i < (int)kFirstPer; ++i) s_out[first_pad(t * kFirstPer + i)] = min(v[i], right);
  __syncthreads();
  for (uint32_t h = h0 + t; h < h1; h += 256) {
    const uint32_t f = s_out[first_pad(h - h0)];
    first[h] = f;
    if (h == nb - 1) *dLive = f;
  }
  if (h1 == nb && t == 0) first[nb] = n;
}

// sortedPT[j] = src[val[j]] (x, y, z, type bits) for live j.
__global__ __launch_bounds__(256) void k_gather(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                                const float4* __restrict__ src, const uint32_t* __restrict__ dN,
                                                uint32_t deadKey, float4* __restrict__ sortedPT) {
  const uint32_t j = blockIdx.x * 256 + threadIdx.x;
  if (j >= *dN) return;
  if (keys[j] >= deadKey) return;
  sortedPT[j] = src[vals[j]];
}

// ---------------------------------------------------------------------------
// Candidate search (collideD/collideCell, particles_kernel.cu:316-411,
// :443-497): for each live sorted particle j, the nearest particle k != j in
// the 27 neighbouring buckets with a bimolecular reaction for the type pair
// and d^2 <= bindRadius^2 (ties: lower k).  Minimum image on periodic axes.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int replica_of_bucket(const uint32_t* __restrict__ bucketStart, int R, uint32_t key) {
  int lo = 0, hi = R;  // bucketStart[R] = deadKey; find last r with bucketStart[r] <= key
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (bucketStart[mid] <= key) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Candidate test: a cheap type-mask and largest-radius filter first, the
// reaction table only for pairs already within the largest binding radius.
__device__ __forceinline__ void consider(const ReplicaTables& T, const float4 q, uint32_t k, uint32_t j, float px,
                                         float py, float pz, int tj, uint32_t mask, float maxB2, int& best,
                                         float& bestD2) {
  // Sorted records are live particles (types 1 to kMaxTypes - 1; no partner
  // mask has a bit for type 0) and may carry their index above the type bits:
  // the bit extract (v_bfe_u32) reads only the low five bits of its offset.
  const uint32_t bits = __float_as_uint(q.w);
  if (k == j || !__builtin_amdgcn_ubfe(mask, bits, 1u)) return;
  const int tk = (int)(bits & kTypeMask);
  const float dx = min_image(px - q.x, T, 0);
  const float dy = min_image(py - q.y, T, 1);
  const float dz = min_image(pz - q.z, T, 2);
  const float d2 = dot3(dx, dy, dz, dx, dy, dz);
  if (d2 > maxB2) return;
  const int rx = T.pairRxn[tj * kMaxTypes + tk];
  if (d2 <= T.bindR2[rx] && (d2 < bestD2 || (d2 == bestD2 && (int)k < best))) {
    best = (int)k;
    bestD2 = d2;
  }
}

// The whole record is in registers here: without this the compiler defers
// the x component's load into the branch behind the type test, a second
// memory round trip for every candidate of a partner type.
__device__ __forceinline__ void whole(const float4& q) { asm volatile("" ::"v"(q.x), "v"(q.y), "v"(q.z), "v"(q.w)); }

__device__ __forceinline__ void scan_range(const ReplicaTables& T, const float4* __restrict__ spt, uint32_t a,
                                           uint32_t b, uint32_t j, float px, float py, float pz, int tj,
                                           uint32_t mask, float maxB2, int& best, float& bestD2) {
  uint32_t k = a;
  for (; k + 2 <= b; k += 2) {  // two loads in flight
    const float4 q0 = spt[k];
    const float4 q1 = spt[k + 1];
    whole(q0);
    consider(T, q0, k, j, px, py, pz, tj, mask, maxB2, best, bestD2);
    whole(q1);
    consider(T, q1, k + 1, j, px, py, pz, tj, mask, maxB2, best, bestD2);
  }
  if (k < b) {
    const float4 q = spt[k];
    whole(q);
    consider(T, q, k, j, px, py, pz, tj, mask, maxB2, best, bestD2);
  }
}

// Face-aware neighbour search.  Every binding radius is at most the cell
// size, so a partner k of j within the binding radius lies in a neighbouring
// cell only across a face that j is within that radius of: along each axis
// the search takes the cell below only when j is near the lower face, the
// cell above only when it is near the upper face.  This visits a subset of
// the 27 cells that contains every pair within the radius, so the nearest
// partner (and every tie) is the one the full 27-cell scan finds.  The
// near-face margin is widened by 1/64 of a cell, far above the rounding of
// the cell coordinates, so no pair is ever lost to rounding.
constexpr float kFaceMargin = 1.0f / 64.0f;

// Search of one wave.  Lanes hold consecutive sorted particles.  Along x the
// buckets of one grid row have consecutive keys, so the own bucket and the
// x-neighbours a particle needs form one key range [kLo, kHi]: the particles
// of that range that sit in the same wave are found by lane shuffles, the
// parts before or after the wave are read from memory, and each other row
// (y or z neighbour) is one more range of consecutive buckets.  Rows that
// wrap along x, or a binding reach above half a cell, take the per-bucket
// search instead.
__device__ __forceinline__ void wave_search(const ReplicaTables& T, const CandArgs& a, uint32_t j, uint32_t wbase,
                                            uint32_t lane, bool active, uint32_t key, const float4 p, int& best,
                                            bool uniform) {
  const int tj = __float_as_int(p.w) & kTypeMask;
  // one replica in the wave: the per-type entries sit one per lane (lane t
  // holds type t's) and reach each lane by a shuffle, so no load waits for
  // the particle's type; otherwise each lane loads its own replica's entry
  static_assert(kMaxTypes <= 64, "one type per lane");
  const bool typed = active && T.hasBim && tj > 0 && tj < kMaxTypes;
  uint32_t pm;
  float mb;
  if (uniform) {
    const uint32_t pmLane = T.partnerMask[lane % kMaxTypes];
    const float mbLane = T.maxBindR2[lane % kMaxTypes];
    pm = __shfl(pmLane, typed ? tj : 0, 64);
    mb = __shfl(mbLane, typed ? tj : 0, 64);
  } else {
    pm = typed ? T.partnerMask[tj] : 0u;
    mb = typed ? T.maxBindR2[tj] : 0.0f;
  }
  const uint32_t mask = typed ? pm : 0u;
  const float maxB2 = mask ? mb : 0.0f;
  float bestD2 = 3.4e38f;
  const uint32_t nLive = *a.dLive;
  // faces this particle is near
  const float pc[3] = {p.x, p.y, p.z};
  int lo[3], hi[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float q = (pc[d] - T.origin[d]) / T.cellSize[d];
    const float f = q - floorf(q);
    const float rb = sqrtf(maxB2) / T.cellSize[d] + kFaceMargin;
    lo[d] = (mask && f < rb) ? 1 : 0;
    hi[d] = (mask && f > 1.0f - rb) ? 1 : 0;
  }
  const uint32_t local = key - T.bucketOff;
  const uint32_t bx = T.gbits[0], by = T.gbits[1];
  const uint32_t mx = (1u << bx) - 1u, my = (1u << by) - 1u, mz = (1u << T.gbits[2]) - 1u;
  const uint32_t gx = local & mx, gy = (local >> bx) & my, gz = local >> (bx + by);
  const bool rows = mask && !(lo[0] & hi[0]) && !(lo[1] & hi[1]) && !(lo[2] & hi[2]) &&
                    !(lo[0] && gx == 0u) && !(hi[0] && gx == mx);
  const uint32_t kLo = rows ? key - (uint32_t)lo[0] : key;
  const uint32_t kHi = rows ? key + (uint32_t)hi[0] : key;
  // The key range inside the wave.  Keys ascend through the lanes, so each
  // key is a run of consecutive lanes and [kLo, kHi] (at most three
  // consecutive keys) is one stretch of lanes [s, e].  Runs start where the
  // key changes; the stretch is the lane's own run, extended by the run
  // before (after) it when that run holds key kLo (kHi).
  const uint32_t keyPrev = __shfl(key, (int)((lane + 63u) & 63u), 64);
  const uint64_t starts = __ballot(lane == 0u || key != keyPrev);
  const uint64_t upTo = lane == 63u ? ~0ull : ((2ull << lane) - 1ull);  // lanes <= lane
  const int s0 = 63 - __clzll(starts & upTo);  // bit 0 is always set
  const uint64_t after0 = starts & ~upTo;
  const int e0 = after0 ? __ffsll((unsigned long long)after0) - 2 : 63;
  const uint32_t kBefore = __shfl(key, s0 > 0 ? s0 - 1 : 0, 64);
  const uint32_t kAfter = __shfl(key, e0 < 63 ? e0 + 1 : 63, 64);
  int s = s0, e = e0;
  if (kLo < key && s0 > 0 && kBefore == kLo) s = 63 - __clzll(starts & ((1ull << s0) - 1ull));
  if (kHi > key && e0 < 63 && kAfter == kHi) {
    const uint64_t after1 = e0 + 1 == 63 ? 0ull : (starts & ~((2ull << (e0 + 1)) - 1ull));
    e = after1 ? __ffsll((unsigned long long)after1) - 2 : 63;
  }
  // Every other lane of the stretch once, by rotation: at step o the lane
  // reads s + ((lane - s + o) mod len); one exchange per step.  Only lanes
  // with reactive partners keep the loop going.
  const int len = e - s + 1;
  const int myLen = mask ? len : 1;
  for (int o = 1; __any(o < myLen); ++o) {
    const bool take = o < myLen;
    int t = (int)lane - s + o;
    if (t >= len) t -= len;
    const int src = take ? s + t : (int)lane;
    // the neighbour's type (without the index bits) stands in for its w
    const float4 q = make_float4(__shfl(p.x, src, 64), __shfl(p.y, src, 64), __shfl(p.z, src, 64),
                                 __int_as_float(__shfl(tj, src, 64)));
    if (take) consider(T, q, wbase + (uint32_t)src, j, p.x, p.y, p.z, tj, mask, maxB2, best, bestD2);
  }
  // the wave's first and last keys, read while every lane is still active
  const uint32_t key0 = __shfl(key, 0, 64), key63 = __shfl(key, 63, 64);
  if (!mask) return;
  const uint32_t* __restrict__ first = a.first;
  // the parts of the key range before the wave ([first[kLo], wbase), when
  // the wave's first key is in the range) and after it ([wbase + 64,
  // first[kHi + 1]), when its last key is); empty when the range starts or
  // ends inside the wave
  const bool spillLo = key0 >= kLo && wbase > 0;
  const bool spillHi = key63 <= kHi && wbase + 64 < nLive;
  if (rows) {
    // at most one other row per axis: every range bound is loaded first,
    // then the ranges are scanned
    const uint32_t act = (uint32_t)(lo[1] | hi[1]) | ((uint32_t)(lo[2] | hi[2]) << 1);
    const uint32_t ys = (gy + (hi[1] ? 1u : my)) & my;
    const uint32_t zs = (gz + (hi[2] ? 1u : mz)) & mz;
    const uint32_t xa = gx - (uint32_t)lo[0], xb = gx + (uint32_t)hi[0] + 1u;
    const uint32_t oLo = spillLo ? first[kLo] : wbase;
    const uint32_t oHi = spillHi ? first[kHi + 1] : wbase + 64;
    uint32_t fa[3], fb[3];
#pragma unroll
    for (int c = 1; c < 4; ++c) {
      const bool use = ((uint32_t)c & ~act) == 0u;
      const uint32_t yy = (c & 1) ? ys : gy, zz = (c & 2) ? zs : gz;
      const uint32_t row = T.bucketOff + (zz << (bx + by)) + (yy << bx);
      fa[c - 1] = use ? first[row + xa] : 0u;
      fb[c - 1] = use ? first[row + xb] : 0u;
    }
    // the first particle of every range requested together, then the rest
    uint32_t ra[5] = {spillLo ? oLo : 0u, spillHi ? wbase + 64 : 0u, fa[0], fa[1], fa[2]};
    const uint32_t rb[5] = {spillLo ? wbase : 0u, spillHi ? oHi : 0u, fb[0], fb[1], fb[2]};
    float4 q0[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) q0[c] = ra[c] < rb[c] ? a.sortedPT[ra[c]] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      if (ra[c] < rb[c]) {
        consider(T, q0[c], ra[c], j, p.x, p.y, p.z, tj, mask, maxB2, best, bestD2);
        scan_range(T, a.sortedPT, ra[c] + 1, rb[c], j, p.x, p.y, p.z, tj, mask, maxB2, best, bestD2);
      }
    }
    return;
  }
  if (spillLo) scan_range(T, a.sortedPT, first[key], wbase, j, p.x, p.y, p.z, tj, mask, maxB2, best, bestD2);
  if (spillHi) scan_range(T, a.sortedPT, wbase + 64, first[key + 1], j, p.x, p.y, p.z, tj, mask, maxB2, best, bestD2);
  for (int dz = -lo[2]; dz <= hi[2]; ++dz) {
    const uint32_t zz = (gz + (uint32_t)dz) & mz;
    for (int dy = -lo[1]; dy <= hi[1]; ++dy) {
      const uint32_t yy = (gy + (uint32_t)dy) & my;
      const uint32_t row = T.bucketOff + (zz << (bx + by)) + (yy << bx);
      for (int dx = -lo[0]; dx <= hi[0]; ++dx) {
        if (dx == 0 && dy == 0 && dz == 0) continue;  // own bucket done above
        const uint32_t xx = (gx + (uint32_t)dx) & mx;
        scan_range(T, a.sortedPT, first[row + xx], first[row + xx + 1u], j, p.x, p.y, p.z, tj, mask, maxB2, best,
                   bestD2);
      }
    }
  }
}

// The bucket offsets and replica tables arrive as read-only non-aliased
// arguments, so the wave's replica (searched once, for its first key) and the
// table fields load through the scalar cache.
template <int BS>
__global__ __launch_bounds__(BS) void k_candidates(CandArgs a, const uint32_t* __restrict__ bucketStart,
                                                   const int* __restrict__ bucketRep,
                                                   const ReplicaTables* __restrict__ tables) {
  const uint32_t nLive = *a.dLive;
  const uint32_t nBlocks = (nLive + BS - 1u) / BS;
  if (blockIdx.x >= nBlocks) return;
  // neighbouring blocks on one XCD: the cells they both search share its L2
  const uint32_t j = xcd_tile(blockIdx.x, nBlocks) * BS + threadIdx.x;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wbase = j - lane;
  if (wbase >= nLive) return;  // the whole wave is past the live particles
  const bool active = j < nLive;
  // inactive lanes get a key no live particle has
  const uint32_t key = active ? a.keys[j] : 0xFFFFFFFFu;
  const float4 p = active ? a.sortedPT[j] : make_float4(0.f, 0.f, 0.f, 0.f);
  // replica of the wave's first key, then lanes move forward past later
  // replica starts (keys ascend through the wave)
  const uint32_t key0 = __builtin_amdgcn_readfirstlane(key);
  const int rw = bucketRep ? bucketRep[key0 >> a.repShift] : replica_of_bucket(bucketStart, a.R, key0);
  int r = rw;
  if (active)
    while (r + 1 < a.R && bucketStart[r + 1] <= key) ++r;
  int best = -1;
  if (!__any(active && r != rw))
    wave_search(tables[rw], a, j, wbase, lane, active, key, p, best, true);  // one replica: scalar table loads
  else
    wave_search(tables[r], a, j, wbase, lane, active, key, p, best, false);
  if (active) a.cand[j] = best;
}

// ---------------------------------------------------------------------------
// Resolve: thread g owns stream g and walks the sorted particles
// j = s, s + S, ... of its replica.  A pair (j, c) reacts when each is the
// other's candidate; the lower sorted index j draws (coin for prob < 1, then
// theta, phi, chi) and performs bimReact (reaction_kernel.cuh:202-258),
// writing products into the unsorted arrays at the original indices.
// Stream state is loaded only by threads that actually draw.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBS) void k_resolve(ResolveArgs a, const ReplicaTables* __restrict__ tables) {
  __shared__ uint32_t lds_mt[kMtNN * kBS];
  const uint32_t g = blockIdx.x * kBS + threadIdx.x;
  if (g >= a.Stot) return;
  // a wave never spans two replicas (stream ranges are 64-aligned), so the
  // table fields load through the scalar cache
  const int r = __builtin_amdgcn_readfirstlane(a.streamReplica[g]);
  if (r < 0) return;
  const ReplicaTables& T = tables[r];
  if (!T.hasBim) return;
  const uint32_t S = T.numStreams;
  const uint32_t s = g - T.streamOff;
  if (s >= S) return;
  const uint32_t lo = a.first[T.bucketOff];
  const uint32_t hi = a.first[T.bucketOff + T.numBuckets];
  Mt m;
  bool loaded = false;
  uint32_t deaths = 0;
  // candidates of eight consecutive particles of the stream are loaded
  // together; the particles are still handled in increasing j
  constexpr int kAhead = 8;
  int cq[kAhead];
  for (uint32_t j0 = lo + s; j0 < hi; j0 += kAhead * S) {
#pragma unroll
    for (int q = 0; q < kAhead; ++q) {
      const uint32_t jq = j0 + (uint32_t)q * S;
      cq[q] = jq < hi ? a.cand[jq] : -1;
    }
#pragma unroll
    for (int q = 0; q < kAhead; ++q) {
      const uint32_t j = j0 + (uint32_t)q * S;
      const int c = cq[q];
      if (c <= (int)j) continue;
      if (a.cand[c] != (int)j) continue;
      const float4 p1 = a.sortedPT[j];
      const float4 p2 = a.sortedPT[c];
      const int t1 = __float_as_int(p1.w) & kTypeMask, t2 = __float_as_int(p2.w) & kTypeMask;
      const int rx = T.pairRxn[t1 * kMaxTypes + t2];
      if (!loaded) {
        mt_load<kBS>(m, &lds_mt[threadIdx.x], a.mtState, a.mtPar, a.Stot, g);
        loaded = true;
      }
      const float pr = T.prob[rx];
      if (!(pr == 1.0f || rand_real(mt_draw<kBS>(m)) < pr)) continue;
      const float dc1 = T.D[t1], dc2 = T.D[t2];
      const float x = (dc1 == 0.0f && dc2 == 0.0f) ? 0.5f : dc2 / (dc1 + dc2);
      // pos2 taken at its image nearest pos1 on periodic axes
      const float q0 = nearest_image(p1.x, p2.x, T, 0);
      const float q1 = nearest_image(p1.y, p2.y, T, 1);
      const float q2 = nearest_image(p1.z, p2.z, T, 2);
      const float omx = 1.0f - x;
      float rp[3] = {fmaf(x, p1.x, omx * q0), fmaf(x, p1.y, omx * q1), fmaf(x, p1.z, omx * q2)};
      const float th = smd_theta(rand_real(mt_draw<kBS>(m)));
      const float ph = rand_real(mt_draw<kBS>(m)) * 6.28318548f;
      const float ch = rand_real(mt_draw<kBS>(m)) * 6.28318548f;
      float v1[3], v2[3];
      rotate_offsets(th, ph, ch, T.prodPos[rx], v1, v2);
      const uint32_t mol1 = a.packed ? unpack_index(p1.w) : a.vals[j];
      const uint32_t mol2 = a.packed ? unpack_index(p2.w) : a.vals[c];
      float n1[3] = {rp[0] + v1[0], rp[1] + v1[1], rp[2] + v1[2]};
      float n2[3] = {rp[0] + v2[0], rp[1] + v2[1], rp[2] + v2[2]};
  #pragma unroll
      for (int d = 0; d < 3; ++d) {
        if (T.periodic[d]) {
          if (n1[d] >= T.hi[d]) n1[d] -= T.L[d];
          else if (n1[d] < T.lo[d]) n1[d] += T.L[d];
          if (n2[d] >= T.hi[d]) n2[d] -= T.L[d];
          else if (n2[d] < T.lo[d]) n2[d] += T.L[d];
        }
      }
      const int nt1 = T.prodX[rx], nt2 = T.prodY[rx];
      a.pt[mol1] = make_float4(n1[0], n1[1], n1[2], __int_as_float(nt1));
      a.pt[mol2] = make_float4(n2[0], n2[1], n2[2], __int_as_float(nt2));
      a.w[mol1] = 0.0f;
      a.w[mol2] = 0.0f;
      if (nt1 <= 0) {
        ++deaths;
        atomicSub(&a.tileAlive[mol1 / kCompactTile], 1u);
      }
      if (nt2 <= 0) {
        ++deaths;
        atomicSub(&a.tileAlive[mol2 / kCompactTile], 1u);
      }
    }
  }
  if (loaded) mt_store<kBS>(m, a.mtState, a.Stot, g);
  if (deaths) atomicAdd(&a.repDeaths[r], deaths);
}

// ---------------------------------------------------------------------------
// Finalize (one workgroup of 1024): per-replica survivor/birth counts and new
// offsets, and exclusive prefixes of the compaction tile counters (which are
// reset for the next step).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t block_exscan1024(uint32_t v, uint32_t* wsum, uint32_t* total) {
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  uint32_t pre = 0;
  for (uint32_t k = 0; k < w; ++k) pre += wsum[k];
  if (threadIdx.x == 1023) *total = pre + x;
  __syncthreads();
  return pre + x - v;
}

// Exclusive scans of K values per thread over the 1024-thread workgroup in
// one pass (the K shuffle chains run side by side); every thread also gets
// the K totals.
template <int K>
__device__ __forceinline__ void block_exscan1024_k(const uint32_t (&v)[K], uint32_t (&ex)[K], uint32_t (&tot)[K],
                                                   uint32_t (*wsum)[16]) {
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  uint32_t x[K];
#pragma unroll
  for (int k = 0; k < K; ++k) x[k] = v[k];
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const uint32_t y = __shfl_up(x[k], o, 64);
      if (lane >= (uint32_t)o) x[k] += y;
    }
  }
  if (lane == 63) {
#pragma unroll
    for (int k = 0; k < K; ++k) wsum[k][w] = x[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    uint32_t pre = 0, all = 0;
#pragma unroll
    for (uint32_t q = 0; q < 16u; ++q) {
      const uint32_t ws = wsum[k][q];
      pre += q < w ? ws : 0u;
      all += ws;
    }
    ex[k] = pre + x[k] - v[k];
    tot[k] = all;
  }
  __syncthreads();
}

// Compaction tiles per finalize workgroup (1024 threads x 4 consecutive tiles).
constexpr uint32_t kFinTPT = 1;
constexpr uint32_t kFinBlockTiles = 1024 * kFinTPT;

// Per-step bookkeeping after the pair step.  Workgroup g scans the compaction
// tiles [g * 1024, (g + 1) * 1024): exclusive prefixes of both counters inside
// the block, the counters reset, the block's totals published for the
// scatter (which adds the totals of the earlier blocks).  Workgroup 0 also
// lays out every replica's new segment and publishes the step's counts.
__global__ __launch_bounds__(1024) void k_finalize(FinalizeArgs a) {
  __shared__ uint32_t wsum[3][16];
  const uint32_t t0 = blockIdx.x * kFinBlockTiles;
  const uint32_t e0 = t0 + threadIdx.x * kFinTPT;
  // Every load that does not depend on the particle count is issued before
  // that count arrives: this block's compaction-tile counters (up to the
  // launch's tile bound; entries past the count are masked below) and the
  // step counters.  All loads come before any store (the arguments may alias).
  uint32_t va[kFinTPT], vb[kFinTPT];
#pragma unroll
  for (uint32_t k = 0; k < kFinTPT; ++k) {
    const uint32_t e = e0 + k;
    va[k] = e < a.maxTiles ? a.tileAlive[e] : 0u;
    vb[k] = e < a.maxTiles ? a.tileBirth[e] : 0u;
  }
  uint32_t stepPrev = 0;
  unsigned long long psPrev = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    stepPrev = *a.stepCount;
    psPrev = *a.particleSteps;
  }
  const uint32_t n = *a.dN;
  const uint32_t nTiles = (n + kCompactTile - 1) / kCompactTile;
  if (blockIdx.x > 0 && t0 >= nTiles) return;
#pragma unroll
  for (uint32_t k = 0; k < kFinTPT; ++k) {
    if (e0 + k >= nTiles) {
      va[k] = 0u;
      vb[k] = 0u;
    }
  }
  if (blockIdx.x == 0) {
    // replicas
    uint32_t cN = 0, cA = 0, cB = 0;
    for (uint32_t b = 0; b < (uint32_t)a.R; b += 1024) {
      const uint32_t r = b + threadIdx.x;
      uint32_t alive = 0, births = 0, zc = 0, inj = 0;
      if (r < (uint32_t)a.R) {
        alive = a.repCount[r] - a.repDeaths[r];
        births = a.repBirths[r];
        if (a.zCount) zc = a.zCount[r];
        if (a.injCount) inj = a.injCount[r];
      }
      // new segment of replica r: [survivors, births, created, admitted]
      const uint32_t v[3] = {alive + births + zc + inj, alive, births};
      uint32_t ex[3], tot[3];
      block_exscan1024_k<3>(v, ex, tot, wsum);
      if (r < (uint32_t)a.R) {
        a.newOff[r] = cN + ex[0];
        a.newCount[r] = v[0];
        a.aliveBase[r] = cA + ex[1];
        a.birthBase[r] = cB + ex[2];
        a.repAlive[r] = alive;
        a.repDeaths[r] = 0;
        a.repBirths[r] = 0;
        if (a.zCount) a.zeroBase[r] = alive + births;
        if (a.injCount) {
          a.injBase[r] = alive + births + zc;
          a.injCount[r] = 0;
        }
        if (a.retire) a.retire[r] = 0;
      }
      cN += tot[0];
      cA += tot[1];
      cB += tot[2];
    }
    if (threadIdx.x == 0) {
      a.newOff[a.R] = cN;
      *a.dNnext = cN;
      *a.particleSteps = psPrev + n;
      // the host reads the particle count after each step from this ring,
      // tagged with the step number (no copy or event in the stream)
      const uint32_t st = stepPrev + 1u;
      *a.stepCount = st;
      __hip_atomic_store(&a.countRing[st & a.ringMask], ((unsigned long long)st << 32) | cN, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
  // this block's compaction tiles: exclusive prefixes inside the block,
  // counters reset, block totals for the scatter
  uint32_t sum[2] = {0u, 0u};
#pragma unroll
  for (uint32_t k = 0; k < kFinTPT; ++k) {
    const uint32_t e = e0 + k;
    sum[0] += va[k];
    sum[1] += vb[k];
    if (e < nTiles) a.tileAlive[e] = 0;  // the birth counters are cleared by the scatter
  }
  uint32_t ex[2], tot[2];
  block_exscan1024_k<2>(sum, ex, tot, wsum);
  uint32_t ra = ex[0], rb = ex[1];
#pragma unroll
  for (uint32_t k = 0; k < kFinTPT; ++k) {
    const uint32_t e = e0 + k;
    if (e < nTiles) {
      a.tileAlivePre[e] = ra;
      a.tileBirthPre[e] = rb;
    }
    ra += va[k];
    rb += vb[k];
  }
  if (threadIdx.x == 0) {
    a.blockSum[blockIdx.x] = tot[0];
    a.blockSum[a.blockStride + blockIdx.x] = tot[1];
  }
}

// ---------------------------------------------------------------------------
// Scatter: order-preserving compaction (the intent of compactPosAndTypeD,
// particles_kernel.cu:97-123, done out of place).  Replica r's new segment is
// [survivors of r in index order, births of r in parent index order].
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_scatter(ScatterArgs a, const uint32_t* __restrict__ repOff) {
  __shared__ uint32_t s_wa[16], s_wb[16];
  const uint32_t n = *a.dN;
  const uint32_t tile = blockIdx.x;
  if (tile * kCompactTile >= n) return;
  const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
  const uint64_t lt = lanemask_lt();
  constexpr int kItems = kCompactTile / 256;  // chunks of 64 per wave
  constexpr uint32_t kWaveSpan = kCompactTile / 4;  // consecutive elements per wave
  bool al[kItems], bi[kItems];
  uint32_t ra[kItems], rb[kItems];
  float4 pv[kItems];
  float wv[kItems];
  int bt[kItems];
  uint32_t ca = 0, cb = 0;
  // every load of the survivors' data before the first store (a wait for a
  // load also waits for the stores issued before it)
  const uint32_t tileBirths = a.tileBirth[tile];  // births the producer counted here
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const uint32_t e = tile * kCompactTile + w * kWaveSpan + it * 64 + lane;
    const bool v = e < n;
    pv[it] = v ? a.pt[e] : make_float4(0.f, 0.f, 0.f, 0.f);
    wv[it] = v ? a.w[e] : 0.0f;
  }
  // birth types are read only in tiles with births (zero everywhere else)
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const uint32_t e = tile * kCompactTile + w * kWaveSpan + it * 64 + lane;
    bt[it] = (tileBirths != 0u && e < n) ? a.birthType[e] : 0;
  }
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const uint32_t e = tile * kCompactTile + w * kWaveSpan + it * 64 + lane;
    const bool v = e < n;
    al[it] = v && __float_as_int(pv[it].w) > 0;
    bi[it] = v && bt[it] > 0;
    const uint64_t am = __ballot(al[it]);
    const uint64_t bm = __ballot(bi[it]);
    ra[it] = ca + (uint32_t)__popcll(am & lt);
    rb[it] = cb + (uint32_t)__popcll(bm & lt);
    ca += (uint32_t)__popcll(am);
    cb += (uint32_t)__popcll(bm);
  }
  uint32_t pa = a.tileAlivePre[tile], pb = a.tileBirthPre[tile];
  // totals of the finalize blocks before this tile's block
  {
    const uint32_t g = tile / kFinBlockTiles;
    uint32_t sa = 0, sb = 0;
    for (uint32_t q = lane; q < g; q += 64) {
      sa += a.blockSum[q];
      sb += a.blockSum[a.blockStride + q];
    }
    pa += wave_sum(sa);
    pb += wave_sum(sb);
  }
  if (lane == 0) {
    s_wa[w] = ca;
    s_wb[w] = cb;
  }
  __syncthreads();
  for (uint32_t k = 0; k < w; ++k) {
    pa += s_wa[k];
    pb += s_wb[k];
  }
  // replica of the wave's first element (last r with repOff[r] <= e), by a
  // wave-uniform search through the scalar cache; each lane's elements only
  // move forward through the replicas from there
  int r = 0;
  {
    const uint32_t e0 = tile * kCompactTile + w * kWaveSpan;
    int lo = 0, hi = a.R;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (repOff[mid] <= e0) lo = mid;
      else hi = mid;
    }
    r = lo;
  }
  // the replica's offsets in registers, reloaded only when the walk moves on
  uint32_t nOff = a.newOff[r], aBase = a.aliveBase[r], bBase = a.birthBase[r], rAlive = a.repAlive[r];
  uint32_t nextOff = r + 1 < a.R ? repOff[r + 1] : 0xFFFFFFFFu;
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const uint32_t e = tile * kCompactTile + w * kWaveSpan + it * 64 + lane;
    if (!(al[it] || bi[it])) continue;
    while (nextOff <= e) {
      ++r;
      nOff = a.newOff[r];
      aBase = a.aliveBase[r];
      bBase = a.birthBase[r];
      rAlive = a.repAlive[r];
      nextOff = r + 1 < a.R ? repOff[r + 1] : 0xFFFFFFFFu;
    }
    if (al[it]) {
      const uint32_t dst = nOff + (pa + ra[it] - aBase);
      a.ptOut[dst] = pv[it];
      a.wOut[dst] = wv[it];
    }
    if (bi[it]) {
      const uint32_t dst = nOff + rAlive + (pb + rb[it] - bBase);
      const float4 q = a.addedPos[e];
      a.ptOut[dst] = make_float4(q.x, q.y, q.z, __int_as_float(bt[it]));
      a.wOut[dst] = q.w;
      a.birthType[e] = 0;  // taken: zero again for the next step
    }
  }
  if (threadIdx.x == 0 && tileBirths != 0u) a.tileBirth[tile] = 0;
}

// ---------------------------------------------------------------------------
// Serving: admitted particles enter their replica's segment after its births;
// stream ranges of an admitted replica are (re)initialised; a finished
// replica's particles are removed by the producer.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_inject(InjectArgs a) {
  const uint32_t q = blockIdx.x * 256 + threadIdx.x;
  if (q >= a.nInj) return;
  const int r = a.injRep[q];
  const uint32_t dst = a.newOff[r] + a.injBase[r] + (q - a.injSrcOff[r]);
  const float4 p = a.injPos[q];
  a.ptOut[dst] = make_float4(p.x, p.y, p.z, __int_as_float(a.injType[q]));
  a.wOut[dst] = p.w;
}

// ---------------------------------------------------------------------------
// Order-0 reactions (reaction_cpu.cpp:36-63, run on the host by the reference
// before each step and appended after the births, particleSystem.cpp:833-837,
// 989-991).  One thread per replica draws, from the replica's own creation
// stream, for each order-0 reaction in index order: the Poisson count
// (stats.cpp:poisrandF multiplication method: multiply closed-interval
// uniforms until the product drops to exp(-mean)), then for each new
// particle x, y, z = lo + u (hi - lo) with u on the open interval
// (particleSystem.h:randPos, stats.h:unirandOOF), w = 1.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float real_closed(uint32_t v) { return (float)((double)v * (1.0 / 4294967295.0)); }
__device__ __forceinline__ float real_open(uint32_t v) { return (float)(((double)v + 0.5) * (1.0 / 4294967296.0)); }

__global__ __launch_bounds__(64) void k_zero(ZeroArgs a) {
  __shared__ uint32_t lds_mt[kMtNN * 64];
  const int r = blockIdx.x * 64 + threadIdx.x;
  if (r >= a.R) return;
  const ReplicaTables& T = a.tables[r];
  if (T.numZero == 0) {
    a.zCount[r] = 0u;
    return;
  }
  Mt m;
  mt_load<64>(m, &lds_mt[threadIdx.x], a.zMt, a.zPar, (uint32_t)a.R, (uint32_t)r);
  const uint32_t cap = T.zeroMax;
  uint32_t cnt = 0;
  bool over = false;
  for (int q = 0; q < T.numZero && !over; ++q) {
    const int rx = T.zeroList[q];
    const float g = T.zeroG[q];
    uint32_t nmol = 0;
    for (int p = 0; p < T.zeroPieces[q] && nmol <= cap; ++p) {
      // bounded: a count above cap is an error anyway
      for (float t = real_closed(mt_draw<64>(m)); t > g && nmol <= cap; t *= real_closed(mt_draw<64>(m))) ++nmol;
    }
    const int type = T.prodX[rx];
    for (uint32_t j = 0; j < nmol; ++j) {
      if (cnt >= cap) {
        over = true;
        break;
      }
      float x[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) x[d] = real_open(mt_draw<64>(m)) * T.L[d] + T.lo[d];
      a.zPos[T.zeroOff + cnt] = make_float4(x[0], x[1], x[2], 1.0f);
      a.zType[T.zeroOff + cnt] = type;
      ++cnt;
    }
  }
  if (over) atomicOr(a.status, kStatusZeroOverflow);
  a.zCount[r] = cnt;
  mt_store<64>(m, a.zMt, (uint32_t)a.R, (uint32_t)r);
}

__global__ __launch_bounds__(64) void k_zero_place(const ReplicaTables* __restrict__ tables,
                                                   const float4* __restrict__ zPos, const int* __restrict__ zType,
                                                   const uint32_t* __restrict__ zCount,
                                                   const uint32_t* __restrict__ newOff,
                                                   const uint32_t* __restrict__ zeroBase, float4* __restrict__ ptOut,
                                                   float* __restrict__ wOut) {
  const int r = blockIdx.x;
  const uint32_t c = zCount[r];
  if (c == 0) return;
  const uint32_t src = tables[r].zeroOff, dst = newOff[r] + zeroBase[r];
  for (uint32_t j = threadIdx.x; j < c; j += 64) {
    const float4 p = zPos[src + j];
    ptOut[dst + j] = make_float4(p.x, p.y, p.z, __int_as_float(zType[src + j]));
    wOut[dst + j] = p.w;
  }
}

__global__ __launch_bounds__(256) void k_unpack(const float4* __restrict__ pt, const float* __restrict__ w, uint32_t n,
                                                float4* __restrict__ pos4, int* __restrict__ type) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float4 p = pt[i];
  pos4[i] = make_float4(p.x, p.y, p.z, w[i]);
  type[i] = __float_as_int(p.w);
}

__global__ __launch_bounds__(256) void k_mt_init_range(uint32_t* __restrict__ st, const uint32_t* __restrict__ seeds,
                                                       uint32_t Stot, uint32_t off, uint32_t count) {
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= count) return;
  const uint32_t g = off + k;
  uint32_t v = seeds[k];
  st[g] = v;
  for (int i = 1; i < kMtNN; ++i) {
    v = 1812433253u * (v ^ (v >> 30)) + (uint32_t)i;
    st[(size_t)i * Stot + g] = v;
  }
  st[(size_t)kMtNN * Stot + g] = 0u;
}

// ---------------------------------------------------------------------------
// Segmented sort of an ensemble's keys.  The canonical particle order is
// replica-major: replica r holds [segOff[r], segOff[r] + segCount[r]), and
// every key there is bucketOff[r] + (a cell of r's grid, below 2^localBits)
// or deadKey.  Sorting each segment on its local cell index (deadKey taken as
// all ones) and moving every segment's dead particles behind all live ones
// gives exactly the stable sort of all keys by global key: live particles
// replica by replica in cell order, then the removed ones in index order.  The
// local index of grids up to 2^17 buckets (plus the removed mark) fits two
// digits of at most 9 bits where the global key needs three passes (24 bits
// for c4's 512 grids).  Tiles never straddle two segments.  Each pass:
// per-tile digit counts (k_seg_up, one row of digits per tile; the first pass
// also counts each segment's removed particles), per segment and digit an
// exclusive scan over the segment's tiles (k_seg_scan; in the first pass one
// extra workgroup lays out where each segment's live and removed particles
// end up), then ranking and scatter (k_seg_down, the body of k_sort_down; each
// workgroup turns its segment's digit totals into digit starts itself).
// ---------------------------------------------------------------------------
constexpr uint32_t kSegDeadLocal = 0xFFFFFFFFu;
constexpr uint32_t kSegHead = 0x80000000u;  // tile table: first tile of its segment
constexpr uint32_t kSegScanLong = 64;  // segments up to this many tiles: one scan workgroup each

__device__ __forceinline__ uint32_t seg_local(uint32_t key, uint32_t bucketOff, uint32_t deadKey) {
  return key == deadKey ? kSegDeadLocal : key - bucketOff;
}

// Tile table: tile t covers keys [tpos[t].x, tpos[t].x + tpos[t].y) of
// segment tseg[t] & ~kSegHead (kSegHead set on a segment's first tile), and
// segment r's tiles start at segFirst[r]; *nLong counts the segments of more
// than kSegScanLong tiles.  One workgroup over all segments, which also clears
// the per-segment removed counts.  Tiles stay inside the particle count and the table inside its
// capacity; a segment layout that disagrees with either sets the sort-layout
// status bit.
// Exclusive scan of one value per thread over BS threads (BS / 64 waves,
// wsum one word per wave); total gets the sum of all of them.
template <int BS>
__device__ __forceinline__ uint32_t block_exscan_bs(uint32_t v, uint32_t* wsum, uint32_t& total) {
  constexpr uint32_t NW = BS / 64;
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  uint32_t pre = 0, all = 0;
#pragma unroll
  for (uint32_t q = 0; q < NW; ++q) {
    const uint32_t ws = wsum[q];
    pre += q < w ? ws : 0u;
    all += ws;
  }
  __syncthreads();
  total = all;
  return pre + x - v;
}

template <int BS>
__device__ __forceinline__ void seg_tiles_body(const SegTilesArgs& a, uint32_t* wsum, uint32_t* sLong) {
  if (threadIdx.x == 0) *sLong = 0;
  __syncthreads();
  const uint32_t nAll = *a.dN;
  bool bad = false;
  uint32_t carry = 0;
  for (uint32_t b = 0; b < (uint32_t)a.R; b += BS) {
    const uint32_t r = b + threadIdx.x;
    const uint32_t cnt = r < (uint32_t)a.R ? a.segCount[r] : 0u;
    const uint32_t v = (cnt + a.tileSize - 1u) / a.tileSize;
    uint32_t tot;
    const uint32_t ex = block_exscan_bs<BS>(v, wsum, tot);
    if (r < (uint32_t)a.R) {
      a.repDead[r] = 0u;
      a.segFirst[r] = carry + ex;
      if (v > kSegScanLong) atomicAdd(sLong, 1u);
      const uint32_t off = a.segOff[r];
      for (uint32_t k = 0; k < v; ++k) {
        const uint32_t idx = carry + ex + k;
        const uint32_t start = off + k * a.tileSize;
        uint32_t len = min(a.tileSize, cnt - k * a.tileSize);
        if (start >= nAll || len > nAll - start) {
          bad = true;
          len = start < nAll ? nAll - start : 0u;
        }
        if (idx < a.maxTiles) {
          a.tpos[idx] = make_uint2(start, len);
          a.tseg[idx] = r | (k == 0 ? kSegHead : 0u);
        } else {
          bad = true;
        }
      }
    }
    carry += tot;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    *a.nTiles = min(carry, a.maxTiles);
    *a.nLong = *sLong;
  }
  if (bad && a.status) atomicOr(a.status, kStatusSortLayout);
}

__global__ __launch_bounds__(1024) void k_seg_tiles(SegTilesArgs a) {
  __shared__ uint32_t wsum[16];
  __shared__ uint32_t s_long;
  seg_tiles_body<1024>(a, wsum, &s_long);
}

template <int BITS, int ITEMS, bool COUNT_DEAD>
__global__ __launch_bounds__(256) void k_seg_up(const uint32_t* __restrict__ keys, const uint2* __restrict__ tpos,
                                                const uint32_t* __restrict__ tseg,
                                                const uint32_t* __restrict__ nTiles,
                                                const uint32_t* __restrict__ bucketOff, uint32_t deadKey,
                                                uint32_t* __restrict__ counts, int shift,
                                                uint32_t* __restrict__ repDead) {
  constexpr int RADIX = 1 << BITS;
  __shared__ uint32_t hist[RADIX];
  __shared__ uint32_t s_dead;
  const uint32_t T = *nTiles;
  if (blockIdx.x >= T) return;
  const uint32_t tile = xcd_tile(blockIdx.x, T);
  const uint2 ps = tpos[tile];
  const uint32_t seg = tseg[tile] & ~kSegHead;
  const uint32_t bo = bucketOff[seg];
  uint32_t kk[ITEMS];
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const uint32_t e = k * 256 + threadIdx.x;
    kk[k] = e < ps.y ? keys[ps.x + e] : 0u;
  }
  for (int d = threadIdx.x; d < RADIX; d += 256) hist[d] = 0;
  if (COUNT_DEAD && threadIdx.x == 0) s_dead = 0;
  __syncthreads();
  uint32_t dead = 0;
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const bool valid = k * 256 + threadIdx.x < ps.y;
    const uint32_t d = (seg_local(kk[k], bo, deadKey) >> shift) & (RADIX - 1);
    const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
    const uint64_t vm = __ballot(valid);
    if (__all(d == d0 || !valid)) {
      if ((threadIdx.x & 63u) == 0u && vm) atomicAdd(&hist[d0], (uint32_t)__popcll(vm));
    } else if (valid) {
      atomicAdd(&hist[d], 1u);
    }
    if (COUNT_DEAD) dead += (uint32_t)__popcll(__ballot(valid && kk[k] == deadKey));
  }
  if (COUNT_DEAD && (threadIdx.x & 63u) == 0u && dead) atomicAdd(&s_dead, dead);
  __syncthreads();
  for (int d = threadIdx.x; d < RADIX; d += 256) counts[(size_t)tile * RADIX + d] = hist[d];
  if (COUNT_DEAD && threadIdx.x == 0 && s_dead) atomicAdd(&repDead[seg], s_dead);
}

// Where each segment's particles end up after the last pass: live ones
// replica by replica from 0 (liveStart), removed ones after all live ones in
// replica order (nLive + deadStart).  Run by one 256-thread workgroup.
__device__ void seg_layout(const uint32_t* __restrict__ segCount, const uint32_t* __restrict__ repDead, int R,
                           uint32_t* __restrict__ liveStart, uint32_t* __restrict__ deadStart,
                           uint32_t* __restrict__ nLive, uint32_t* wsum, uint32_t* tot) {
  uint32_t cA = 0, cD = 0;
  for (uint32_t b = 0; b < (uint32_t)R; b += 256) {
    const uint32_t r = b + threadIdx.x;
    const uint32_t dead = r < (uint32_t)R ? repDead[r] : 0u;
    const uint32_t alive = r < (uint32_t)R ? segCount[r] - dead : 0u;
    const uint32_t exA = block_exscan256(alive, wsum, tot);
    const uint32_t totA = *tot;
    const uint32_t exD = block_exscan256(dead, wsum, tot);
    const uint32_t totD = *tot;
    if (r < (uint32_t)R) {
      liveStart[r] = cA + exA;
      deadStart[r] = cD + exD;
    }
    cA += totA;
    cD += totD;
    __syncthreads();  // *tot is rewritten by the next chunk
  }
  if (threadIdx.x == 0) *nLive = cA;
}

// Exclusive scans of the digit counts over each segment's tiles: counts[t][d]
// becomes the sum over the segment's tiles before t, segTot[seg][d] the
// segment's total.  Workgroup r < R scans segment r when it has at most
// kSegScanLong tiles: a thread per digit walks the tiles in order, every row
// read whole by the workgroup.  Workgroup R + d scans digit d of the longer
// segments (none in most ensembles: then it returns at once) by block scans.
// With `layout`, workgroup R + RADIX lays out the segments' live and removed
// particles.
template <int BITS>
__global__ __launch_bounds__(256) void k_seg_scan(uint32_t* __restrict__ counts,
                                                  const uint32_t* __restrict__ segFirst,
                                                  const uint32_t* __restrict__ segCount, int R, uint32_t tileSize,
                                                  const uint32_t* __restrict__ nTiles,
                                                  const uint32_t* __restrict__ nLong, uint32_t* __restrict__ segTot,
                                                  int layout, const uint32_t* __restrict__ repDead,
                                                  uint32_t* __restrict__ liveStart, uint32_t* __restrict__ deadStart,
                                                  uint32_t* __restrict__ nLive) {
  constexpr int RADIX = 1 << BITS;
  constexpr int DPT = (RADIX + 255) / 256;
  constexpr int BATCH = 16;  // tile rows requested together
  constexpr int PER = 8;     // tiles per thread in a long segment's block scan
  __shared__ uint32_t wsum[4];
  __shared__ uint32_t s_tot;
  const uint32_t g = blockIdx.x;
  const uint32_t T = *nTiles;
  if (g < (uint32_t)R) {
    const uint32_t nt = (segCount[g] + tileSize - 1u) / tileSize;
    const uint32_t t0 = segFirst[g];
    // a table that does not fit was reported by k_seg_tiles
    if (nt == 0u || nt > kSegScanLong || t0 >= T || nt > T - t0) return;
#pragma unroll
    for (int q = 0; q < DPT; ++q) {
      const uint32_t d = threadIdx.x + 256u * q;
      if (d >= (uint32_t)RADIX) break;
      uint32_t* col = counts + (size_t)t0 * RADIX + d;
      uint32_t run = 0;
      for (uint32_t k0 = 0; k0 < nt; k0 += BATCH) {
        uint32_t v[BATCH];
#pragma unroll
        for (int i = 0; i < BATCH; ++i) v[i] = k0 + i < nt ? col[(size_t)(k0 + i) * RADIX] : 0u;
#pragma unroll
        for (int i = 0; i < BATCH; ++i) {
          if (k0 + i < nt) col[(size_t)(k0 + i) * RADIX] = run;
          run += v[i];
        }
      }
      segTot[(size_t)g * RADIX + d] = run;
    }
    return;
  }
  if (g < (uint32_t)R + RADIX) {
    if (*nLong == 0u) return;
    const uint32_t d = g - (uint32_t)R;
    for (int r = 0; r < R; ++r) {
      const uint32_t nt = (segCount[r] + tileSize - 1u) / tileSize;
      const uint32_t t0 = segFirst[r];
      if (nt <= kSegScanLong || t0 >= T || nt > T - t0) continue;
      uint32_t* col = counts + (size_t)t0 * RADIX + d;
      uint32_t carry = 0;
      for (uint32_t b = 0; b < nt; b += 256u * PER) {
        const uint32_t e0 = b + threadIdx.x * PER;
        uint32_t v[PER], sum = 0;
#pragma unroll
        for (int i = 0; i < PER; ++i) {
          v[i] = e0 + i < nt ? col[(size_t)(e0 + i) * RADIX] : 0u;
          sum += v[i];
        }
        uint32_t run = carry + block_exscan256(sum, wsum, &s_tot);
#pragma unroll
        for (int i = 0; i < PER; ++i) {
          if (e0 + i < nt) col[(size_t)(e0 + i) * RADIX] = run;
          run += v[i];
        }
        carry += s_tot;
        __syncthreads();  // s_tot is rewritten by the next chunk
      }
      if (threadIdx.x == 0) segTot[(size_t)r * RADIX + d] = carry;
    }
    return;
  }
  if (layout) seg_layout(segCount, repDead, R, liveStart, deadStart, nLive, wsum, &s_tot);
}

template <int BITS, int ITEMS, bool GATHER, bool LAST>
__global__ __launch_bounds__(256) void k_seg_down(const uint32_t* __restrict__ keysIn,
                                                  const uint32_t* __restrict__ valsIn,
                                                  uint32_t* __restrict__ keysOut, uint32_t* __restrict__ valsOut,
                                                  const uint32_t* __restrict__ counts,
                                                  const uint2* __restrict__ tpos,
                                                  const uint32_t* __restrict__ tseg,
                                                  const uint32_t* __restrict__ nTiles,
                                                  const uint32_t* __restrict__ bucketOff, uint32_t deadKey,
                                                  const uint32_t* __restrict__ segOff,
                                                  const uint32_t* __restrict__ segTot,
                                                  const uint32_t* __restrict__ liveStart,
                                                  const uint32_t* __restrict__ deadStart,
                                                  const uint32_t* __restrict__ nLive, int shift, int identityVals,
                                                  const float4* __restrict__ gatherSrc,
                                                  float4* __restrict__ gatherDst, const uint32_t* __restrict__ dN,
                                                  uint32_t* __restrict__ status, int packIndex) {
  constexpr int RADIX = 1 << BITS;
  constexpr int DPT = (RADIX + 255) / 256;
  constexpr uint32_t TILE = 256u * ITEMS;
  __shared__ uint32_t s_base[RADIX];
  __shared__ uint32_t s_wcnt[4][RADIX];
  __shared__ uint32_t s_tstart[RADIX];
  __shared__ uint32_t s_key[TILE];
  __shared__ uint32_t s_val[TILE];
  __shared__ uint32_t wsum[4];
  __shared__ uint32_t tot;

  const uint32_t T = *nTiles;
  if (blockIdx.x >= T) return;
  const uint32_t tile = xcd_tile(blockIdx.x, T);
  const uint2 ps = tpos[tile];
  const uint32_t seg = tseg[tile] & ~kSegHead, len = ps.y;
  const uint32_t bo = bucketOff[seg];
  // every destination is checked against the particle count: a layout that
  // disagrees with it is reported (sticky status bit), never written
  const uint32_t nAll = *dN;
  bool bad = false;
  const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
  const uint32_t base = w * (64 * ITEMS);  // tile-local
  uint32_t kk[ITEMS], vv[ITEMS], rk[ITEMS];
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    const uint32_t e = base + it * 64 + lane;
    const bool valid = e < len;
    kk[it] = valid ? keysIn[ps.x + e] : 0u;
    vv[it] = identityVals ? ps.x + e : (valid ? valsIn[ps.x + e] : 0u);
  }
  // where each digit of this tile starts: the segment's digit start (its
  // digit totals scanned here; live and removed particles apart in the last
  // pass) + the earlier tiles of the segment
  {
    const uint32_t* st = segTot + (size_t)seg * RADIX;
    uint32_t tv[DPT], sum = 0;
#pragma unroll
    for (int q = 0; q < DPT; ++q) {
      const uint32_t d = t * DPT + q;
      tv[q] = d < (uint32_t)RADIX ? st[d] : 0u;
      sum += tv[q];
    }
    uint32_t run = block_exscan256(sum, wsum, &tot);
    uint32_t liveBase = 0, deadBase = 0;
    if (LAST) {
      liveBase = liveStart[seg];
      deadBase = *nLive + deadStart[seg];
    } else {
      liveBase = segOff[seg];
    }
#pragma unroll
    for (int q = 0; q < DPT; ++q) {
      const uint32_t d = t * DPT + q;
      if (d < (uint32_t)RADIX) {
        // removed particles, and only they, have the top digit in the last pass
        const uint32_t b0 = (LAST && d == (uint32_t)RADIX - 1u) ? deadBase : liveBase + run;
        s_base[d] = b0 + counts[(size_t)tile * RADIX + d];
      }
      run += tv[q];
    }
  }
  for (uint32_t d = t; d < (uint32_t)RADIX; d += 256) {
    s_wcnt[0][d] = 0;
    s_wcnt[1][d] = 0;
    s_wcnt[2][d] = 0;
    s_wcnt[3][d] = 0;
  }
  __syncthreads();

  uint32_t dg[ITEMS];
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    const uint32_t e = base + it * 64 + lane;
    const bool valid = e < len;
    const uint32_t d = (seg_local(kk[it], bo, deadKey) >> shift) & (RADIX - 1);
    dg[it] = d;
    const uint64_t mm = wave_digit_peers<BITS>(d, valid);
    // every lane of the digit writes the same new count (a lane past the
    // end counts its valid peers only, which is the same number)
    const uint32_t prev = s_wcnt[w][d];
    rk[it] = prev + lanes_below(mm);
    s_wcnt[w][d] = prev + (uint32_t)__popcll(mm);
  }
  __syncthreads();
  {
    uint32_t c[DPT][4], tsum = 0;
#pragma unroll
    for (int q = 0; q < DPT; ++q) {
      const uint32_t d = t * DPT + q;
#pragma unroll
      for (int k = 0; k < 4; ++k) c[q][k] = (d < (uint32_t)RADIX) ? s_wcnt[k][d] : 0u;
      tsum += c[q][0] + c[q][1] + c[q][2] + c[q][3];
    }
    uint32_t run = block_exscan256(tsum, wsum, &tot);
#pragma unroll
    for (int q = 0; q < DPT; ++q) {
      const uint32_t d = t * DPT + q;
      if (d < (uint32_t)RADIX) {
        s_wcnt[0][d] = 0;
        s_wcnt[1][d] = c[q][0];
        s_wcnt[2][d] = c[q][0] + c[q][1];
        s_wcnt[3][d] = c[q][0] + c[q][1] + c[q][2];
        s_tstart[d] = run;
      }
      run += c[q][0] + c[q][1] + c[q][2] + c[q][3];
    }
    __syncthreads();
  }
#pragma unroll
  for (int it = 0; it < ITEMS; ++it) {
    const uint32_t e = base + it * 64 + lane;
    if (e < len) {
      const uint32_t d = dg[it];
      const uint32_t lp = s_tstart[d] + s_wcnt[w][d] + rk[it];
      s_key[lp] = kk[it];
      s_val[lp] = vv[it];
    }
  }
  __syncthreads();
  if (GATHER) {
    uint32_t dk[ITEMS], dv[ITEMS], dd[ITEMS];
    float4 rec[ITEMS];
#pragma unroll
    for (int q = 0; q < ITEMS; ++q) {
      const uint32_t idx = t + q * 256u;
      const bool ok = idx < len;
      const uint32_t k = ok ? s_key[idx] : deadKey;
      const uint32_t d = (seg_local(k, bo, deadKey) >> shift) & (RADIX - 1);
      dk[q] = k;
      dv[q] = ok ? s_val[idx] : 0u;
      dd[q] = ok ? s_base[d] + (idx - s_tstart[d]) : 0u;
      rec[q] = (ok && k < deadKey) ? gatherSrc[dv[q]] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int q = 0; q < ITEMS; ++q) {
      if (t + q * 256u < len) {
        if (dd[q] >= nAll) {
          bad = true;
          continue;
        }
        keysOut[dd[q]] = dk[q];
        if (packIndex) {
          // the record carries its particle's index above the type bits
          if (dk[q] < deadKey)
            gatherDst[dd[q]] = make_float4(rec[q].x, rec[q].y, rec[q].z,
                                           __int_as_float(pack_index_type(dv[q], __float_as_int(rec[q].w))));
        } else {
          valsOut[dd[q]] = dv[q];
          if (dk[q] < deadKey) gatherDst[dd[q]] = rec[q];
        }
      }
    }
  } else {
    for (uint32_t idx = t; idx < len; idx += 256) {
      const uint32_t k = s_key[idx];
      const uint32_t d = (seg_local(k, bo, deadKey) >> shift) & (RADIX - 1);
      const uint32_t dst = s_base[d] + (idx - s_tstart[d]);
      if (dst >= nAll) {
        bad = true;
        continue;
      }
      keysOut[dst] = k;
      valsOut[dst] = s_val[idx];
    }
  }
  if (bad && status) atomicOr(status, kStatusSortLayout);
}

// ---------------------------------------------------------------------------
// Launch wrappers.
// ---------------------------------------------------------------------------
static inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

hipError_t launch_mt_init(uint32_t* st, const uint32_t* seeds, uint32_t S, hipStream_t stream) {
  if (S == 0) return hipSuccess;
  hipLaunchKernelGGL(k_mt_init, dim3(cdiv(S, 256)), dim3(256), 0, stream, st, seeds, S);
  return hipGetLastError();
}

hipError_t launch_mt_fill(uint32_t* st, const uint32_t* par, uint32_t S, uint32_t nStreams, int count, int batch3,
                          uint32_t* out, hipStream_t stream) {
  if (nStreams == 0) return hipSuccess;
  hipLaunchKernelGGL(k_mt_fill, dim3(cdiv(nStreams, kBS)), dim3(kBS), 0, stream, st, par, S, nStreams, count,
                     batch3, out);
  return hipGetLastError();
}

hipError_t launch_producer(const ProducerArgs& a, int blockSize, hipStream_t stream) {
  // the stream of global id g is g's in every shape: the results do not
  // depend on the block size; one more workgroup for the tile table
  const unsigned extra = a.segTilesOn ? 1u : 0u;
  if (blockSize == 512)
    hipLaunchKernelGGL(k_producer<512>, dim3(cdiv(a.Stot, 512) + extra), dim3(512), 0, stream, a, a.tables);
  else if (blockSize == 128)
    hipLaunchKernelGGL(k_producer<128>, dim3(cdiv(a.Stot, 128) + extra), dim3(128), 0, stream, a, a.tables);
  else
    hipLaunchKernelGGL(k_producer<256>, dim3(cdiv(a.Stot, 256) + extra), dim3(256), 0, stream, a, a.tables);
  return hipGetLastError();
}

int sort_passes(uint32_t keyBits, int radixBits) { return (int)((keyBits + radixBits - 1) / radixBits); }

size_t sort_counts_words(uint32_t maxN, int radixBits) {
  // enough for both layouts: per-tile counters of the three-kernel sort, or
  // histograms + tile counters + error flag + per-pass status words
  const size_t radix = (size_t)1 << radixBits;
  const size_t tiles = cdiv(maxN, kSortTile) + 1;
  const size_t passes = (size_t)((32 + radixBits - 1) / radixBits);
  const size_t a = radix * (cdiv(maxN, 2048) + 1) + radix;  // three-kernel sort, tiles down to 2048 keys
  const size_t b = passes * radix + passes + 4 + passes * tiles * radix;
  return a > b ? a : b;
}

hipError_t launch_sort_3k(SortBuffers& b, const uint32_t* dN, uint32_t maxN, uint32_t keyBits, int radixBits,
                          bool identityVals, hipStream_t stream);

static hipError_t launch_sort_onesweep(SortBuffers& b, const uint32_t* dN, uint32_t maxN, uint32_t keyBits,
                                       int radixBits, bool identityVals, hipStream_t stream) {
  const uint32_t maxTiles = cdiv(maxN, kSortTile);
  if (maxTiles == 0) return hipSuccess;
  const int radix = 1 << radixBits;
  const int passes = sort_passes(keyBits, radixBits);
  uint32_t* hist = b.counts;
  uint32_t* counters = hist + (size_t)passes * radix;
  // layout inside b.counts: hist | tile counters | 4 words | per-pass status
  // words; a look-back overrun is flagged in the caller's sticky status word
  // when there is one, else in the first of the 4 words
  uint32_t* status = counters + passes + 4;
  uint32_t* err = b.status ? b.status : counters + passes;
  const uint32_t zeroWords = (uint32_t)((size_t)passes * radix + passes + 4 + (size_t)passes * maxTiles * radix);
  hipLaunchKernelGGL(k_os_zero, dim3(std::min<unsigned>(cdiv(zeroWords, 256), 1024)), dim3(256), 0, stream, b.counts,
                     zeroWords);
  const uint32_t histBlocks = std::min<uint32_t>(cdiv(maxN, 16384), 1024);
  if (radixBits == 8)
    hipLaunchKernelGGL(k_os_hist<8>, dim3(histBlocks), dim3(256), 0, stream, b.keys, dN, maxN, passes, hist);
  else if (radixBits == 4)
    hipLaunchKernelGGL(k_os_hist<4>, dim3(histBlocks), dim3(256), 0, stream, b.keys, dN, maxN, passes, hist);
  else
    return hipErrorInvalidValue;
  for (int p = 0; p < passes; ++p) {
    const int shift = p * radixBits;
    const int idv = (identityVals && p == 0) ? 1 : 0;
    uint32_t* st = status + (size_t)p * maxTiles * radix;
    if (radixBits == 8)
      hipLaunchKernelGGL(k_os_pass<8>, dim3(maxTiles), dim3(256), 0, stream, b.keys, b.vals, b.keysAlt, b.valsAlt, dN,
                         maxN, hist + (size_t)p * radix, st, counters + p, err, shift, idv);
    else
      hipLaunchKernelGGL(k_os_pass<4>, dim3(maxTiles), dim3(256), 0, stream, b.keys, b.vals, b.keysAlt, b.valsAlt, dN,
                         maxN, hist + (size_t)p * radix, st, counters + p, err, shift, idv);
    std::swap(b.keys, b.keysAlt);
    std::swap(b.vals, b.valsAlt);
  }
  if (b.gatherDst)
    hipLaunchKernelGGL(k_gather, dim3(maxTiles * (kSortTile / 256)), dim3(256), 0, stream, b.keys, b.vals,
                       b.gatherSrc, dN, b.gatherDeadKey, b.gatherDst);
  return hipGetLastError();
}

// Segmented sort: scratch layout (words) and launch.  Both passes share one
// tile table (tiles of 4096 keys inside segments); the counters are sized
// for the wider (second) digit.
constexpr int kSegItems = 16;

static uint32_t seg_max_tiles(uint32_t maxN, int R) { return cdiv(maxN, 256u * kSegItems) + (uint32_t)R; }

size_t seg_sort_words(uint32_t maxN, int R, int maxBits) {
  const size_t radix = (size_t)1 << maxBits;
  const size_t mt = seg_max_tiles(maxN, R);
  return 3 * mt + 64 + radix * (mt + 1) + (size_t)R * radix + 4 * (size_t)R + 64;
}

void seg_sort_digits(uint32_t maxGridBits, int& lowBits, int& highBits) {
  // removed particles need one value above every live cell.  10-bit digits
  // (replica grids of 2^18 buckets and more) made the sort slower than the
  // three 8-bit passes over global keys (serving layout: 59 against 47 us),
  // so those grids take the global sort.
  const uint32_t need = maxGridBits + 1u;
  static const int opts[3][2] = {{8, 8}, {9, 8}, {9, 9}};
  lowBits = highBits = 0;
  for (const auto& o : opts)
    if ((uint32_t)(o[0] + o[1]) >= need) {
      lowBits = o[0];
      highBits = o[1];
      return;
    }
}

namespace {
struct SegScratch {
  uint2* tpos;
  uint32_t *tseg, *nTiles, *nLive, *nLong, *counts, *segTot, *repDead, *liveStart, *deadStart, *segFirst;
  uint32_t maxTiles;
};
}  // namespace

// One digit pass of the segmented sort (BITS wide at `shift`).  The first
// pass also counts removed particles and lays out the segments; the last
// places removed particles after every replica's live ones and gathers the
// records when asked.
template <int BITS, bool FIRST, bool LAST>
static void seg_pass(SortBuffers& b, const SegSortArgs& g, const SegScratch& s, const uint32_t* dN, int shift,
                     bool identityVals, hipStream_t stream) {
  constexpr int RADIX = 1 << BITS;
  const dim3 grid(s.maxTiles), blk(256);
  hipLaunchKernelGGL((k_seg_up<BITS, kSegItems, FIRST>), grid, blk, 0, stream, b.keys, s.tpos, s.tseg, s.nTiles,
                     g.bucketOff, g.deadKey, s.counts, shift, s.repDead);
  hipLaunchKernelGGL((k_seg_scan<BITS>), dim3((uint32_t)g.R + RADIX + (FIRST ? 1u : 0u)), blk, 0, stream, s.counts,
                     s.segFirst, g.segCount, g.R, 256u * kSegItems, s.nTiles, s.nLong, s.segTot, FIRST ? 1 : 0,
                     s.repDead, s.liveStart, s.deadStart, s.nLive);
  const int idv = identityVals ? 1 : 0;
  if (!LAST)
    hipLaunchKernelGGL((k_seg_down<BITS, kSegItems, false, false>), grid, blk, 0, stream, b.keys, b.vals, b.keysAlt,
                       b.valsAlt, s.counts, s.tpos, s.tseg, s.nTiles, g.bucketOff, g.deadKey, g.segOff,
                       s.segTot, s.liveStart, s.deadStart, s.nLive, shift, idv, nullptr, nullptr, dN, b.status, 0);
  else if (b.gatherDst)
    hipLaunchKernelGGL((k_seg_down<BITS, kSegItems, true, true>), grid, blk, 0, stream, b.keys, b.vals, b.keysAlt,
                       b.valsAlt, s.counts, s.tpos, s.tseg, s.nTiles, g.bucketOff, g.deadKey, g.segOff,
                       s.segTot, s.liveStart, s.deadStart, s.nLive, shift, idv, b.gatherSrc, b.gatherDst, dN,
                       b.status, b.packIndex);
  else
    hipLaunchKernelGGL((k_seg_down<BITS, kSegItems, false, true>), grid, blk, 0, stream, b.keys, b.vals, b.keysAlt,
                       b.valsAlt, s.counts, s.tpos, s.tseg, s.nTiles, g.bucketOff, g.deadKey, g.segOff,
                       s.segTot, s.liveStart, s.deadStart, s.nLive, shift, idv, nullptr, nullptr, dN, b.status, 0);
}

// Where each table lives in g.scratch (seg_sort_words words).
static SegScratch seg_scratch(const SegSortArgs& g, uint32_t maxN) {
  SegScratch s;
  s.maxTiles = seg_max_tiles(maxN, g.R);
  uint32_t* w = g.scratch;
  s.tpos = (uint2*)w;
  w += 2 * (size_t)s.maxTiles;
  s.tseg = w;
  w += s.maxTiles;
  s.nTiles = w;
  s.nLive = w + 1;
  s.nLong = w + 2;
  w += 64;
  const int mb = g.lowBits > g.highBits ? g.lowBits : g.highBits;
  s.counts = w;
  w += ((size_t)1 << mb) * (s.maxTiles + 1);
  s.segTot = w;
  w += (size_t)g.R << mb;
  s.repDead = w;
  s.liveStart = w + g.R;
  s.deadStart = w + 2 * (size_t)g.R;
  s.segFirst = w + 3 * (size_t)g.R;
  return s;
}

SegTilesArgs seg_tiles_args(const SegSortArgs& g, const uint32_t* dN, uint32_t maxN, uint32_t* status) {
  const SegScratch s = seg_scratch(g, maxN);
  SegTilesArgs t;
  t.segOff = g.segOff;
  t.segCount = g.segCount;
  t.R = g.R;
  t.tileSize = 256u * kSegItems;
  t.maxTiles = s.maxTiles;
  t.tpos = s.tpos;
  t.tseg = s.tseg;
  t.nTiles = s.nTiles;
  t.segFirst = s.segFirst;
  t.nLong = s.nLong;
  t.repDead = s.repDead;
  t.dN = dN;
  t.status = status;
  return t;
}

template <int LB, int HB>
static hipError_t seg_sort_run(SortBuffers& b, const SegSortArgs& g, const uint32_t* dN, uint32_t maxN,
                               bool identityVals, hipStream_t stream) {
  const SegScratch s = seg_scratch(g, maxN);
  if (!g.tilesLaid)
    hipLaunchKernelGGL(k_seg_tiles, dim3(1), dim3(1024), 0, stream, seg_tiles_args(g, dN, maxN, b.status));
  seg_pass<LB, true, false>(b, g, s, dN, 0, identityVals, stream);
  std::swap(b.keys, b.keysAlt);
  std::swap(b.vals, b.valsAlt);
  seg_pass<HB, false, true>(b, g, s, dN, LB, false, stream);
  std::swap(b.keys, b.keysAlt);
  std::swap(b.vals, b.valsAlt);
  return hipGetLastError();
}

hipError_t launch_sort_seg(SortBuffers& b, const SegSortArgs& g, const uint32_t* dN, uint32_t maxN,
                           bool identityVals, hipStream_t stream) {
  if (maxN == 0 || g.R <= 0) return hipSuccess;
  // digit widths 7 to 10 bits in either pass
#define SMD_SEG_HI(LO)                                                        \
  switch (g.highBits) {                                                      \
    case 7: return seg_sort_run<LO, 7>(b, g, dN, maxN, identityVals, stream);  \
    case 8: return seg_sort_run<LO, 8>(b, g, dN, maxN, identityVals, stream);  \
    case 9: return seg_sort_run<LO, 9>(b, g, dN, maxN, identityVals, stream);  \
    case 10: return seg_sort_run<LO, 10>(b, g, dN, maxN, identityVals, stream); \
    default: return hipErrorInvalidValue;                                    \
  }
  switch (g.lowBits) {
    case 7: SMD_SEG_HI(7)
    case 8: SMD_SEG_HI(8)
    case 9: SMD_SEG_HI(9)
    case 10: SMD_SEG_HI(10)
    default: return hipErrorInvalidValue;
  }
#undef SMD_SEG_HI
}

hipError_t launch_sort(SortBuffers& b, const uint32_t* dN, uint32_t maxN, uint32_t keyBits, int radixBits,
                       bool identityVals, hipStream_t stream) {
  // the single sweep keeps every digit's histogram in LDS at once: 4- and 8-bit digits only
  if (b.algo == 1 || (radixBits != 4 && radixBits != 8))
    return launch_sort_3k(b, dN, maxN, keyBits, radixBits, identityVals, stream);
  return launch_sort_onesweep(b, dN, maxN, keyBits, radixBits, identityVals, stream);
}

// One digit pass of the count / scan / scatter sort.
template <int BITS, int ITEMS>
static hipError_t sort3k_pass(SortBuffers& b, const uint32_t* dN, uint32_t maxTiles, uint32_t stride,
                              uint32_t* totals, int shift, bool idv, bool gather, hipStream_t stream) {
  hipLaunchKernelGGL((k_sort_up<BITS, ITEMS>), dim3(maxTiles), dim3(256), 0, stream, b.keys, dN, b.counts, stride,
                     shift);
  // measured: one wave per digit is faster up to ~100 tiles (16k keys: 1.5 us
  // per step), slower at 512 (1M keys: 1 us per digit pass)
  if (maxTiles <= 128)
    hipLaunchKernelGGL((k_sort_rowscan_wave<BITS, ITEMS>), dim3(((1 << BITS) + 3) / 4), dim3(256), 0, stream,
                       b.counts, totals, dN, stride);
  else
    hipLaunchKernelGGL((k_sort_rowscan<BITS, ITEMS>), dim3(1 << BITS), dim3(256), 0, stream, b.counts, totals, dN,
                       stride);
  if (gather)
    hipLaunchKernelGGL((k_sort_down<BITS, ITEMS, true>), dim3(maxTiles), dim3(256), 0, stream, b.keys, b.vals,
                       b.keysAlt, b.valsAlt, b.counts, totals, dN, stride, shift, idv ? 1 : 0, b.gatherSrc,
                       b.gatherDst, b.gatherDeadKey, b.packIndex);
  else
    hipLaunchKernelGGL((k_sort_down<BITS, ITEMS, false>), dim3(maxTiles), dim3(256), 0, stream, b.keys, b.vals,
                       b.keysAlt, b.valsAlt, b.counts, totals, dN, stride, shift, idv ? 1 : 0, nullptr, nullptr, 0u,
                       0);
  return hipGetLastError();
}

// 2048-key tiles up to 6M keys, 4096-key tiles beyond (measured: 1M-4M keys
// sort 10-20% faster with the smaller tiles, 9M keys 3% slower).  The
// expected count decides when the caller gives one (the bound maxN can be
// several times the live count); the grid always covers maxN.
static int sort3k_items(uint32_t maxN, uint32_t sizeHint) {
  const uint32_t expect = sizeHint ? std::min(sizeHint, maxN) : maxN;
  return expect <= 6000000u ? 8 : 16;
}

size_t sort3k_counts_words(uint32_t maxN, uint32_t sizeHint, int radixBits) {
  const uint32_t maxTiles = cdiv(maxN, 256u * sort3k_items(maxN, sizeHint));
  return ((size_t)1 << radixBits) * (maxTiles + 1) + ((size_t)1 << radixBits);
}

hipError_t launch_sort_3k(SortBuffers& b, const uint32_t* dN, uint32_t maxN, uint32_t keyBits, int radixBits,
                          bool identityVals, hipStream_t stream) {
  const int items = sort3k_items(maxN, b.sizeHint);
  const uint32_t maxTiles = cdiv(maxN, 256u * items);
  if (maxTiles == 0) return hipSuccess;
  const uint32_t stride = maxTiles + 1;
  const int radix = 1 << radixBits;
  uint32_t* totals = b.counts + (size_t)radix * stride;
  const int passes = sort_passes(keyBits, radixBits);
  for (int p = 0; p < passes; ++p) {
    const int shift = p * radixBits;
    const bool idv = identityVals && p == 0;
    const bool gather = b.gatherDst != nullptr && p == passes - 1;
    hipError_t e = hipErrorInvalidValue;
    if (items == 8) {
      if (radixBits == 8) e = sort3k_pass<8, 8>(b, dN, maxTiles, stride, totals, shift, idv, gather, stream);
      else if (radixBits == 4) e = sort3k_pass<4, 8>(b, dN, maxTiles, stride, totals, shift, idv, gather, stream);
      else if (radixBits == 11) e = sort3k_pass<11, 8>(b, dN, maxTiles, stride, totals, shift, idv, gather, stream);
    } else {
      if (radixBits == 8) e = sort3k_pass<8, 16>(b, dN, maxTiles, stride, totals, shift, idv, gather, stream);
      else if (radixBits == 4) e = sort3k_pass<4, 16>(b, dN, maxTiles, stride, totals, shift, idv, gather, stream);
      else if (radixBits == 11) e = sort3k_pass<11, 16>(b, dN, maxTiles, stride, totals, shift, idv, gather, stream);
    }
    if (e != hipSuccess) return e;
    uint32_t* tk = b.keys;
    b.keys = b.keysAlt;
    b.keysAlt = tk;
    uint32_t* tv = b.vals;
    b.vals = b.valsAlt;
    b.valsAlt = tv;
  }
  return hipGetLastError();
}

size_t first_words(uint32_t nb) { return (size_t)nb + 1; }

hipError_t launch_first(const uint32_t* keys, const uint32_t* dN, uint32_t maxN, uint32_t nb, uint32_t* first,
                        uint32_t* dLive, const ReplicaTables* tables, uint32_t slotBits, uint32_t nSlots,
                        hipStream_t stream) {
  // spans must not straddle slots for the skip test
  if (slotBits && ((1u << slotBits) < kFirstSpan || !tables)) slotBits = 0;
  hipLaunchKernelGGL(k_first, dim3(cdiv(nb, kFirstSpan)), dim3(256), 0, stream, keys, dN, maxN, nb, first, dLive,
                     tables, slotBits, nSlots);
  return hipGetLastError();
}

hipError_t launch_gather(const uint32_t* keys, const uint32_t* vals, const float4* src, const uint32_t* dN,
                         uint32_t maxN, uint32_t deadKey, float4* sortedPT, hipStream_t stream) {
  if (maxN == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather, dim3(cdiv(maxN, 256)), dim3(256), 0, stream, keys, vals, src, dN, deadKey, sortedPT);
  return hipGetLastError();
}

hipError_t launch_candidates(const CandArgs& a, uint32_t maxN, int blockSize, hipStream_t stream) {
  if (maxN == 0) return hipSuccess;
  // one result per sorted particle whatever the shape: waves are always 64
  // consecutive particles
  if (blockSize == 64)
    hipLaunchKernelGGL(k_candidates<64>, dim3(cdiv(maxN, 64)), dim3(64), 0, stream, a, a.bucketStart, a.bucketRep,
                       a.tables);
  else if (blockSize == 128)
    hipLaunchKernelGGL(k_candidates<128>, dim3(cdiv(maxN, 128)), dim3(128), 0, stream, a, a.bucketStart,
                       a.bucketRep, a.tables);
  else
    hipLaunchKernelGGL(k_candidates<256>, dim3(cdiv(maxN, 256)), dim3(256), 0, stream, a, a.bucketStart,
                       a.bucketRep, a.tables);
  return hipGetLastError();
}

hipError_t launch_resolve(const ResolveArgs& a, hipStream_t stream) {
  hipLaunchKernelGGL(k_resolve, dim3(cdiv(a.Stot, kBS)), dim3(kBS), 0, stream, a, a.tables);
  return hipGetLastError();
}

uint32_t finalize_blocks(uint32_t maxN) {
  return std::max<uint32_t>(1u, cdiv(cdiv(maxN, kCompactTile), kFinBlockTiles));
}

hipError_t launch_finalize(const FinalizeArgs& args, uint32_t maxN, hipStream_t stream) {
  const uint32_t g = finalize_blocks(maxN);
  if (g > args.blockStride) return hipErrorInvalidValue;
  FinalizeArgs a = args;
  a.maxTiles = cdiv(maxN, kCompactTile);
  hipLaunchKernelGGL(k_finalize, dim3(g), dim3(1024), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_inject(const InjectArgs& a, hipStream_t stream) {
  if (a.nInj == 0) return hipSuccess;
  hipLaunchKernelGGL(k_inject, dim3(cdiv(a.nInj, 256)), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_zero(const ZeroArgs& a, hipStream_t stream) {
  if (a.R <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero, dim3(cdiv(a.R, 64)), dim3(64), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_zero_place(int R, const ReplicaTables* tables, const float4* zPos, const int* zType,
                             const uint32_t* zCount, const uint32_t* newOff, const uint32_t* zeroBase, float4* ptOut,
                             float* wOut, hipStream_t stream) {
  if (R <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_place, dim3(R), dim3(64), 0, stream, tables, zPos, zType, zCount, newOff, zeroBase,
                     ptOut, wOut);
  return hipGetLastError();
}

hipError_t launch_unpack(const float4* pt, const float* w, uint32_t n, float4* pos4, int* type, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_unpack, dim3(cdiv(n, 256)), dim3(256), 0, stream, pt, w, n, pos4, type);
  return hipGetLastError();
}

hipError_t launch_mt_init_range(uint32_t* st, const uint32_t* seeds, uint32_t Stot, uint32_t off, uint32_t count,
                                hipStream_t stream) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_mt_init_range, dim3(cdiv(count, 256)), dim3(256), 0, stream, st, seeds, Stot, off, count);
  return hipGetLastError();
}


hipError_t launch_scatter(const ScatterArgs& a, uint32_t maxN, hipStream_t stream) {
  if (maxN == 0) return hipSuccess;
  hipLaunchKernelGGL(k_scatter, dim3(cdiv(maxN, kCompactTile)), dim3(256), 0, stream, a, a.repOff);
  return hipGetLastError();
}

}  // namespace smd
