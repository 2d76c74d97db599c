// Host orchestration of the timestep over an ensemble of replicas (a plain
// simulation is an ensemble of one).  Replaces ParticleSystem::_initialize
// (particleSystem.cpp:573-751) and ParticleSystem::update (:829-1001): the
// particle state stays in HBM, particle counts live on the device, and the
// host only reads a count back (asynchronously) to bound grid sizes and to
// grow the buffers before they could overflow.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "smd_engine.h"

namespace smd {

#define HIPCHK(x)                                                                          \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) {                                                                \
      set_error(std::string("HIP error '") + hipGetErrorString(e_) + "' in " #x);          \
      return SMD_ERR_HIP;                                                                  \
    }                                                                                      \
  } while (0)

template <typename T>
static hipError_t dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  return hipMalloc((void**)p, count * sizeof(T));
}

template <typename T>
static void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

static inline uint32_t round64(uint32_t x) { return (x + 63u) & ~63u; }

// Caller layout (pos4 = x, y, z, w; int types) to and from the canonical
// device layout (pt = x, y, z, type bits; w apart).
static void pack_particles(const float* pos4, const int* types, size_t n, std::vector<float4>& pt,
                           std::vector<float>& w) {
  pt.resize(n);
  w.resize(n);
  for (size_t i = 0; i < n; ++i) {
    float tb;
    std::memcpy(&tb, &types[i], sizeof(float));
    pt[i] = make_float4(pos4[4 * i], pos4[4 * i + 1], pos4[4 * i + 2], tb);
    w[i] = pos4[4 * i + 3];
  }
}

static hipError_t upload_particles(const float* pos4, const int* types, size_t n, float4* dpt, float* dw) {
  std::vector<float4> pt;
  std::vector<float> w;
  pack_particles(pos4, types, n, pt, w);
  hipError_t e = hipMemcpy(dpt, pt.data(), n * sizeof(float4), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dw, w.data(), n * sizeof(float), hipMemcpyHostToDevice);
  return e;
}

static hipError_t download_particles(const float4* dpt, const float* dw, size_t n, float* pos4, int* types) {
  std::vector<float4> pt(n);
  std::vector<float> w(pos4 ? n : 0);
  hipError_t e = hipMemcpy(pt.data(), dpt, n * sizeof(float4), hipMemcpyDeviceToHost);
  if (e == hipSuccess && pos4) e = hipMemcpy(w.data(), dw, n * sizeof(float), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return e;
  for (size_t i = 0; i < n; ++i) {
    if (pos4) {
      pos4[4 * i] = pt[i].x;
      pos4[4 * i + 1] = pt[i].y;
      pos4[4 * i + 2] = pt[i].z;
      pos4[4 * i + 3] = w[i];
    }
    if (types) std::memcpy(&types[i], &pt[i].w, sizeof(int));
  }
  return hipSuccess;
}

Engine::Engine(int dev, int rbits) : device(dev), radixBits(rbits == 4 || rbits == 8 || rbits == 11 ? rbits : 0) {}

Engine::~Engine() {
  drop_graph();
  if (stream) {
    (void)hipSetDevice(device);
    (void)hipStreamSynchronize(stream);
  }
  free_particle_buffers();
  dfree(dTables);
  dfree(dStreamReplica);
  dfree(dBucketStart);
  dfree(dBucketRep);
  dfree(dRepOff);
  dfree(dRepCount);
  dfree(dNewOff);
  dfree(dNewCount);
  dfree(dAliveBase);
  dfree(dBirthBase);
  dfree(dRepAlive);
  dfree(dRepDeaths);
  dfree(dRepBirths);
  dfree(dCounters);
  dfree(dParticleSteps);
  dfree(dMtState);
  dfree(dMtPar);
  dfree(dGauss);
  dfree(dFirst);
  dfree(dRetire);
  dfree(dInjCount);
  dfree(dInjBase);
  dfree(dStatus);
  dfree(dZMt);
  dfree(dZPar);
  dfree(dZCount);
  dfree(dZeroBase);
  dfree(dZPos);
  dfree(dZType);
  dfree(dInjSrcOff);
  dfree(dInjPos);
  dfree(dInjType);
  dfree(dInjRep);
  if (countRing) (void)hipHostFree(countRing);
  dfree(dStepCount);
  if (stream) (void)hipStreamDestroy(stream);
}

void Engine::free_particle_buffers() {
  for (int k = 0; k < 2; ++k) {
    dfree(dPT[k]);
    dfree(dW[k]);
  }
  dfree(dAddedPos);
  dfree(dBirthType);
  dfree(dSortedPT);
  dfree(dCand);
  dfree(dViewPos);
  dfree(dViewType);
  viewCap = 0;
  dfree(sb.keys);
  dfree(sb.vals);
  dfree(sb.keysAlt);
  dfree(sb.valsAlt);
  dfree(sb.counts);
  dfree(dTileAlive);
  dfree(dTileBirth);
  dfree(dTileAlivePre);
  dfree(dTileBirthPre);
  dfree(dBlockSum);
  dfree(dSegScratch);
}

int Engine::add_replica(const smd_sim_spec& spec, const float* pos4, const int* types, uint32_t n) {
  if (finalized) {
    set_error("add_replica after finalize");
    return -SMD_ERR_INVALID;
  }
  ReplicaTables T;
  std::string err;
  std::vector<smd_rxn_derived> der(spec.num_reactions > 0 ? spec.num_reactions : 1);
  if (!derive_tables(spec, T, der.data(), err)) {
    set_error(err);
    return -SMD_ERR_RXN_SETUP;
  }
  for (uint32_t i = 0; i < n; ++i)
    if (types[i] < 1 || types[i] >= spec.num_types) {
      set_error("initial particle type out of range");
      return -SMD_ERR_INVALID;
    }
  specs.push_back(spec);
  tabs.push_back(T);
  derived.push_back(der);
  pos0.emplace_back(pos4, pos4 + 4 * (size_t)n);
  type0.emplace_back(types, types + n);
  return (int)tabs.size() - 1;
}

smd_status Engine::alloc_particle_buffers(uint32_t c) {
  cap = c;
  for (int k = 0; k < 2; ++k) {
    HIPCHK(dalloc(&dPT[k], cap));
    HIPCHK(dalloc(&dW[k], cap));
  }
  HIPCHK(dalloc(&dAddedPos, cap));
  HIPCHK(dalloc(&dBirthType, cap));
  HIPCHK(dalloc(&dSortedPT, cap));
  HIPCHK(dalloc(&dCand, cap));
  HIPCHK(dalloc(&sb.keys, cap));
  HIPCHK(dalloc(&sb.vals, cap));
  HIPCHK(dalloc(&sb.keysAlt, cap));
  HIPCHK(dalloc(&sb.valsAlt, cap));
  HIPCHK(dalloc(&sb.counts, sort_counts_words(cap, radixBits)));
  nTilesCap = cap / kCompactTile + 2;
  HIPCHK(dalloc(&dTileAlive, nTilesCap));
  HIPCHK(dalloc(&dTileBirth, nTilesCap));
  HIPCHK(dalloc(&dTileAlivePre, nTilesCap));
  HIPCHK(dalloc(&dTileBirthPre, nTilesCap));
  blockStride = finalize_blocks(cap);
  HIPCHK(dalloc(&dBlockSum, 2 * (size_t)blockStride));
  if (segHigh) HIPCHK(dalloc(&dSegScratch, seg_sort_words(cap, (int)R, std::max(segLow, segHigh))));
  HIPCHK(hipMemsetAsync(dTileAlive, 0, nTilesCap * sizeof(uint32_t), stream));
  HIPCHK(hipMemsetAsync(dTileBirth, 0, nTilesCap * sizeof(uint32_t), stream));
  HIPCHK(hipMemsetAsync(dBirthType, 0, (size_t)cap * sizeof(int), stream));
  return SMD_OK;
}

smd_status Engine::finalize(uint32_t capacity) {
  if (finalized) return SMD_OK;
  if (tabs.empty()) {
    set_error("no replicas");
    return SMD_ERR_INVALID;
  }
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  R = (uint32_t)tabs.size();
  // Bucket and stream layout.
  uint64_t buckets = 0, streams = 0, total = 0;
  anySplit = false;
  anyBim = false;
  anyZero = false;
  uint64_t zeros = 0;
  for (uint32_t r = 0; r < R; ++r) {
    ReplicaTables& T = tabs[r];
    T.bucketOff = (uint32_t)buckets;
    buckets += T.numBuckets;
    T.streamOff = (uint32_t)streams;
    streams += round64(T.numStreams);
    total += type0[r].size();
    for (int x = 0; x < T.numRxn; ++x)
      if (specs[r].rxn[x].order == 1 && T.prodY[x] > 0) anySplit = true;
    if (T.hasBim) anyBim = true;
    T.zeroOff = (uint32_t)zeros;
    zeros += T.zeroMax;
    if (T.numZero) anyZero = true;
  }
  if (buckets >= (1ull << 31) || streams >= (1ull << 31) || zeros >= (1ull << 30)) {
    set_error("too many buckets, streams or order-0 creations");
    return SMD_ERR_INVALID;
  }
  zeroTotal = (uint32_t)zeros;
  deadKey = (uint32_t)buckets;
  keyBits = 0;
  while ((1ull << keyBits) <= deadKey) ++keyBits;
  // automatic digit width: 8 bits (11-bit digits save a pass for 17-22-bit
  // keys but each pass costs more than the pass saved: 80 KB of LDS per
  // workgroup and 2048 scan rows)
  if (radixBits == 0) radixBits = 8;
  nb = deadKey + 1;
  // Ensembles sort each replica's segment on its local cells: two digit
  // passes cover the largest replica grid plus the removed mark.
  {
    uint32_t maxGb = 0;
    for (uint32_t r = 0; r < R; ++r) {
      uint32_t b = 0;
      while ((1u << b) < tabs[r].numBuckets) ++b;
      maxGb = std::max(maxGb, b);
    }
    seg_sort_digits(maxGb, segLow, segHigh);
    if (R <= 1) segLow = segHigh = 0;
  }
  Stot = (uint32_t)streams;
  uint64_t c = capacity ? capacity : std::max<uint64_t>(anySplit ? 8 * total : total, 1024);
  if (c < step_bound(total)) c = step_bound(total);
  if (c >= (1ull << 31)) {
    set_error("capacity too large");
    return SMD_ERR_CAPACITY;
  }

  // Per-replica arrays.
  HIPCHK(dalloc(&dTables, R));
  HIPCHK(hipMemcpy(dTables, tabs.data(), R * sizeof(ReplicaTables), hipMemcpyHostToDevice));
  std::vector<int> streamRep(Stot, -1);
  for (uint32_t r = 0; r < R; ++r)
    for (uint32_t k = 0; k < round64(tabs[r].numStreams); ++k) streamRep[tabs[r].streamOff + k] = (int)r;
  HIPCHK(dalloc(&dStreamReplica, Stot));
  HIPCHK(hipMemcpy(dStreamReplica, streamRep.data(), Stot * sizeof(int), hipMemcpyHostToDevice));
  std::vector<uint32_t> bstart(R + 1);
  for (uint32_t r = 0; r < R; ++r) bstart[r] = tabs[r].bucketOff;
  bstart[R] = deadKey;
  HIPCHK(dalloc(&dBucketStart, R + 1));
  HIPCHK(hipMemcpy(dBucketStart, bstart.data(), (R + 1) * sizeof(uint32_t), hipMemcpyHostToDevice));
  // Replica of a bucket by one lookup: every replica owns a power-of-two run
  // of buckets at an offset that is a multiple of the smallest such run, so
  // chunks of that size never straddle two replicas.
  {
    uint32_t minBits = 31;
    for (uint32_t r = 0; r < R; ++r) {
      uint32_t b = 0;
      while ((1u << b) < tabs[r].numBuckets) ++b;
      minBits = std::min(minBits, b);
    }
    const uint64_t chunks = ((uint64_t)deadKey + (1ull << minBits) - 1) >> minBits;
    if (R > 1 && chunks <= (1ull << 22)) {
      std::vector<int> rep((size_t)std::max<uint64_t>(chunks, 1), -1);
      for (uint32_t r = 0; r < R; ++r)
        for (uint64_t c = tabs[r].bucketOff >> minBits; c < ((uint64_t)tabs[r].bucketOff + tabs[r].numBuckets) >> minBits;
             ++c)
          rep[c] = (int)r;
      HIPCHK(dalloc(&dBucketRep, rep.size()));
      HIPCHK(hipMemcpy(dBucketRep, rep.data(), rep.size() * sizeof(int), hipMemcpyHostToDevice));
      repShift = minBits;
    }
  }
  std::vector<uint32_t> off(R + 1), cnt(R);
  uint64_t o = 0;
  for (uint32_t r = 0; r < R; ++r) {
    off[r] = (uint32_t)o;
    cnt[r] = (uint32_t)type0[r].size();
    o += cnt[r];
  }
  off[R] = (uint32_t)o;
  for (uint32_t** p : {&dRepOff, &dNewOff}) HIPCHK(dalloc(p, R + 1));
  for (uint32_t** p : {&dRepCount, &dNewCount, &dAliveBase, &dBirthBase, &dRepAlive, &dRepDeaths, &dRepBirths})
    HIPCHK(dalloc(p, R));
  HIPCHK(hipMemcpy(dRepOff, off.data(), (R + 1) * sizeof(uint32_t), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dRepCount, cnt.data(), R * sizeof(uint32_t), hipMemcpyHostToDevice));
  HIPCHK(hipMemset(dRepDeaths, 0, R * sizeof(uint32_t)));
  HIPCHK(hipMemset(dRepBirths, 0, R * sizeof(uint32_t)));
  HIPCHK(dalloc(&dCounters, 8));
  HIPCHK(dalloc(&dParticleSteps, 1));
  HIPCHK(hipMemset(dParticleSteps, 0, sizeof(unsigned long long)));
  dN = dCounters;
  dNnext = dCounters + 1;
  dLive = dCounters + 2;
  const uint32_t tot32 = (uint32_t)total;
  HIPCHK(hipMemcpy(dN, &tot32, sizeof(uint32_t), hipMemcpyHostToDevice));
  HIPCHK(dalloc(&dFirst, first_words(nb)));

  // Gaussian table: the producer keeps its half t[1 .. kGaussDim / 2], from
  // which the antisymmetry rebuilds every entry (gauss_entry).
  std::vector<float> g(kGaussDim);
  gauss_table(g.data(), kGaussDim);
  for (int u = 0; u < kGaussDim; ++u) {
    const int k = u - 1;
    const float want = (u >= 1 && u <= kGaussDim / 2) ? g[k + 1] : -g[((kGaussDim - u) & (kGaussDim / 2 - 1)) + 1];
    if (std::memcmp(&want, &g[u], sizeof(float)) != 0) {
      set_error("Gaussian table is not antisymmetric");
      return SMD_ERR_INVALID;
    }
  }
  HIPCHK(dalloc(&dGauss, kGaussDim / 2));
  HIPCHK(hipMemcpy(dGauss, g.data() + 1, kGaussDim / 2 * sizeof(float), hipMemcpyHostToDevice));

  // Random streams: stream k of a replica uses MersenneTwister.dat record
  // k % 32768 and seed k of that replica's seed sequence.
  std::string err;
  const uint32_t* dat = mt_dat(err);
  if (!dat) {
    set_error(err);
    return SMD_ERR_IO;
  }
  std::vector<uint32_t> seeds(Stot, 0u), par(3 * (size_t)Stot, 0u);
  for (uint32_t r = 0; r < R; ++r) {
    const ReplicaTables& T = tabs[r];
    if (!stream_seeds(specs[r].seed, T.numStreams, &seeds[T.streamOff], err)) {
      set_error(err);
      return SMD_ERR_IO;
    }
    for (uint32_t k = 0; k < T.numStreams; ++k) {
      const uint32_t rec = k % kMtParamSets;
      par[T.streamOff + k] = dat[4 * rec];
      par[(size_t)Stot + T.streamOff + k] = dat[4 * rec + 1];
      par[2 * (size_t)Stot + T.streamOff + k] = dat[4 * rec + 2];
    }
  }
  HIPCHK(dalloc(&dMtState, (size_t)kMtWords * Stot));
  HIPCHK(dalloc(&dMtPar, 3 * (size_t)Stot));
  HIPCHK(hipMemcpy(dMtPar, par.data(), par.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  uint32_t* dSeeds = nullptr;
  HIPCHK(dalloc(&dSeeds, Stot));
  HIPCHK(hipMemcpy(dSeeds, seeds.data(), Stot * sizeof(uint32_t), hipMemcpyHostToDevice));
  HIPCHK(launch_mt_init(dMtState, dSeeds, Stot, stream));
  HIPCHK(hipStreamSynchronize(stream));
  dfree(dSeeds);
  HIPCHK(dalloc(&dStatus, 1));
  HIPCHK(hipMemset(dStatus, 0, sizeof(uint32_t)));
  sb.status = dStatus;
  if (anyZero) {
    // Replica r's creation stream: the next stream after its S particle
    // streams (seed S of its seed sequence, MersenneTwister.dat record S).
    std::vector<uint32_t> zs(R), zp(3 * (size_t)R);
    for (uint32_t r = 0; r < R; ++r) {
      const uint32_t S = tabs[r].numStreams;
      std::vector<uint32_t> sq(S + 1);
      if (!stream_seeds(specs[r].seed, S + 1, sq.data(), err)) {
        set_error(err);
        return SMD_ERR_IO;
      }
      zs[r] = sq[S];
      const uint32_t rec = S % kMtParamSets;
      for (int w = 0; w < 3; ++w) zp[(size_t)w * R + r] = dat[4 * rec + w];
    }
    HIPCHK(dalloc(&dZMt, (size_t)kMtWords * R));
    HIPCHK(dalloc(&dZPar, 3 * (size_t)R));
    HIPCHK(dalloc(&dZCount, R));
    HIPCHK(dalloc(&dZeroBase, R));
    HIPCHK(dalloc(&dZPos, zeroTotal));
    HIPCHK(dalloc(&dZType, zeroTotal));
    HIPCHK(hipMemcpy(dZPar, zp.data(), zp.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    uint32_t* dzs = nullptr;
    HIPCHK(dalloc(&dzs, R));
    HIPCHK(hipMemcpy(dzs, zs.data(), R * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIPCHK(launch_mt_init(dZMt, dzs, R, stream));
    HIPCHK(hipStreamSynchronize(stream));
    dfree(dzs);
  }

  // Particles.
  smd_status st = alloc_particle_buffers((uint32_t)c);
  if (st != SMD_OK) return st;
  for (uint32_t r = 0; r < R; ++r) {
    if (cnt[r] == 0) continue;
    HIPCHK(upload_particles(pos0[r].data(), type0[r].data(), cnt[r], dPT[0] + off[r], dW[0] + off[r]));
  }
  cur = 0;
  Nub = total;
  knownN = total;
  stepsDone = 0;
  HIPCHK(hipHostMalloc((void**)&countRing, kRing * sizeof(unsigned long long),
                       hipHostMallocMapped | hipHostMallocCoherent));
  std::memset(countRing, 0, kRing * sizeof(unsigned long long));
  HIPCHK(hipHostGetDevicePointer((void**)&dCountRing, countRing, 0));
  HIPCHK(dalloc(&dStepCount, 1));
  HIPCHK(hipMemset(dStepCount, 0, sizeof(uint32_t)));
  lastSeen = 0;
  HIPCHK(hipStreamSynchronize(stream));
  finalized = true;
  return SMD_OK;
}

// Bound on the particle count after all enqueued steps, refreshed from the
// latest count the finalize kernel has published (blocking: after the stream
// has drained, the count after the last step).
void Engine::refresh_bound(bool block) {
  if (lastSeen >= stepsDone) return;
  if (block) (void)hipStreamSynchronize(stream);
  const uint64_t oldest = stepsDone >= kRing ? stepsDone - kRing + 1 : 1;
  for (uint64_t s = stepsDone; s > lastSeen && s >= oldest; --s) {
    const unsigned long long v = __atomic_load_n(&countRing[s & (kRing - 1)], __ATOMIC_ACQUIRE);
    if ((uint32_t)(v >> 32) != (uint32_t)s) continue;
    knownN = (uint32_t)v;
    uint64_t b = knownN;
    for (uint64_t k = 0; k < stepsDone - s && b < (1ull << 40); ++k) b = step_bound(b);
    Nub = std::min<uint64_t>(Nub, b);
    lastSeen = s;
    break;
  }
}

smd_status Engine::grow(uint64_t need) {
  drop_graph();
  HIPCHK(hipStreamSynchronize(stream));
  uint64_t nc = std::max<uint64_t>(need, 2ull * cap);
  if (nc >= (1ull << 31)) {
    set_error("particle count exceeds the 2^31 capacity limit");
    return SMD_ERR_CAPACITY;
  }
  const uint32_t n = (uint32_t)knownN;
  float4* keepPT = nullptr;
  float* keepW = nullptr;
  HIPCHK(dalloc(&keepPT, n));
  HIPCHK(dalloc(&keepW, n));
  if (n) {
    HIPCHK(hipMemcpy(keepPT, dPT[cur], n * sizeof(float4), hipMemcpyDeviceToDevice));
    HIPCHK(hipMemcpy(keepW, dW[cur], n * sizeof(float), hipMemcpyDeviceToDevice));
  }
  free_particle_buffers();
  smd_status st = alloc_particle_buffers((uint32_t)nc);
  if (st != SMD_OK) return st;
  cur = 0;
  if (n) {
    HIPCHK(hipMemcpy(dPT[0], keepPT, n * sizeof(float4), hipMemcpyDeviceToDevice));
    HIPCHK(hipMemcpy(dW[0], keepW, n * sizeof(float), hipMemcpyDeviceToDevice));
  }
  dfree(keepPT);
  dfree(keepW);
  HIPCHK(hipStreamSynchronize(stream));
  return SMD_OK;
}

smd_status Engine::ensure_capacity() {
  refresh_bound(false);
  const uint64_t needAfter = step_bound(Nub);
  if (needAfter <= cap) return SMD_OK;
  refresh_bound(true);
  const uint64_t need2 = step_bound(Nub);
  if (need2 <= cap) return SMD_OK;
  return grow(2 * need2);
}

smd_status Engine::enqueue_step(StageTimer* timer) {
  const uint32_t maxN = (uint32_t)std::min<uint64_t>(Nub, cap);
  smd_status st = enqueue_kernels(maxN, timer);
  if (st != SMD_OK) return st;
  ++stepsDone;
  Nub = std::min<uint64_t>(step_bound(Nub), (uint64_t)cap);
  return SMD_OK;
}

void Engine::drop_graph(int which) {
  const int lo = which < 0 ? 0 : which, hi = which < 0 ? 1 : which;
  bool any = false;
  for (int p = lo; p <= hi; ++p) any = any || gexec[p];
  if (!any) return;
  // A launch of the graph may still be queued or running: the executable
  // graph (and the kernel arguments it owns) must outlive it.
  if (stream) {
    (void)hipSetDevice(device);
    (void)hipStreamSynchronize(stream);
  }
  for (int p = lo; p <= hi; ++p) {
    if (gexec[p]) (void)hipGraphExecDestroy(gexec[p]);
    gexec[p] = nullptr;
    gMaxN[p] = 0;
  }
}

Engine::HostSnapshot Engine::snapshot() const {
  return HostSnapshot{cur, dRepOff, dNewOff, dRepCount, dNewCount, dN, dNnext, sb};
}

void Engine::restore(const HostSnapshot& s) {
  cur = s.cur;
  dRepOff = s.repOff;
  dNewOff = s.newOff;
  dRepCount = s.repCount;
  dNewCount = s.newCount;
  dN = s.n;
  dNnext = s.nnext;
  sb = s.sb;
}

// Two consecutive steps captured as one hipGraph: after two steps every
// ping-pong buffer is back in place, so the graph can be replayed as long as
// the particle bound stays within the grid sizes it was captured with.  An
// eager step (an admission in a slot engine) flips the ping-pong parity, so
// one graph is kept per parity.
smd_status Engine::step_graph(int pairs) {
  for (int k = 0; k < pairs; ++k) {
    refresh_bound(false);
    uint64_t need = step_bound(step_bound(Nub));  // output of the second step
    if (need > cap) {
      refresh_bound(true);
      need = step_bound(step_bound(Nub));
      if (need > cap) {
        smd_status st = grow(2 * need);  // waits for the stream and drops the graphs
        if (st != SMD_OK) return st;
      }
    }
    uint64_t bound = step_bound(Nub);  // input bound of the second step
    const int p = cur;
    if (gexec[p] && bound > gMaxN[p]) {
      // a stale count alone must not force a new capture: read the exact one
      refresh_bound(true);
      bound = step_bound(Nub);
    }
    uint64_t want = bound;
    if (gexec[p] && bound > gMaxN[p]) {
      // outgrown: grids with headroom, so that a growing system re-captures rarely
      want = std::max<uint64_t>(bound, (uint64_t)gMaxN[p] + gMaxN[p] / 2);
      drop_graph(p);
    }
    if (!gexec[p]) {
      const uint32_t maxN = (uint32_t)std::min<uint64_t>(want, cap);
      const HostSnapshot before = snapshot();
      hipGraph_t g = nullptr;
      HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      smd_status st = enqueue_kernels(maxN, nullptr);
      if (st == SMD_OK) st = enqueue_kernels(maxN, nullptr);
      hipError_t e = hipStreamEndCapture(stream, &g);
      // nothing ran: the host-side buffer selection goes back to where it was
      restore(before);
      if (st != SMD_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
      }
      HIPCHK(e);
      hipGraphExec_t x = nullptr;
      e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) {
        if (x) (void)hipGraphExecDestroy(x);
        HIPCHK(e);
      }
      gexec[p] = x;
      gMaxN[p] = maxN;
    }
    HIPCHK(hipGraphLaunch(gexec[p], stream));
    // two steps: every ping-pong selection is back where it was
    stepsDone += 2;
    Nub = std::min<uint64_t>(step_bound(step_bound(Nub)), (uint64_t)cap);
  }
  return SMD_OK;
}

// The step's key sort: segmented by replica when the engine holds several
// replicas and the knob allows it (the same sorted keys, values and records
// either way), else over all keys at once.
bool Engine::seg_sorted() const { return segHigh && segSort && sb.algo == 1; }

SegSortArgs Engine::seg_args() const {
  SegSortArgs g;
  g.R = (int)R;
  g.segOff = dRepOff;
  g.segCount = dRepCount;
  g.bucketOff = dBucketStart;
  g.deadKey = deadKey;
  g.lowBits = segLow;
  g.highBits = segHigh;
  g.scratch = dSegScratch;
  return g;
}

hipError_t Engine::sort_keys(uint32_t maxN, bool tilesLaid) {
  if (seg_sorted()) {
    SegSortArgs g = seg_args();
    g.tilesLaid = tilesLaid ? 1 : 0;
    return launch_sort_seg(sb, g, dN, maxN, true, stream);
  }
  return launch_sort(sb, dN, maxN, keyBits, radixBits, true, stream);
}

smd_status Engine::enqueue_kernels(uint32_t maxN, StageTimer* timer) {
  auto mark = [&](int stage) {
    if (timer) (void)hipEventRecord(timer->ev[stage], stream);
  };
  mark(0);
  ProducerArgs pa;
  pa.tables = dTables;
  pa.streamReplica = dStreamReplica;
  pa.repOff = dRepOff;
  pa.repCount = dRepCount;
  pa.mtState = dMtState;
  pa.mtPar = dMtPar;
  pa.Stot = Stot;
  pa.gauss = dGauss;
  pa.pt = dPT[cur];
  pa.w = dW[cur];
  pa.addedPos = dAddedPos;
  pa.birthType = dBirthType;
  pa.key = sb.keys;
  pa.deadKey = deadKey;
  pa.repDeaths = dRepDeaths;
  pa.repBirths = dRepBirths;
  pa.tileAlive = dTileAlive;
  pa.tileBirth = dTileBirth;
  pa.retire = dRetire;
  // the segmented sort's tile table by one more workgroup of the producer's
  // launch (it reads only the per-replica counts the step starts with)
  const bool tilesLaid = anyBim && seg_sorted();
  if (tilesLaid) {
    pa.segTilesOn = 1;
    pa.segTiles = seg_tiles_args(seg_args(), dN, maxN, sb.status);
  }
  HIPCHK(launch_producer(pa, producerBlock, stream));
  if (anyZero) {
    ZeroArgs za;
    za.tables = dTables;
    za.R = (int)R;
    za.zMt = dZMt;
    za.zPar = dZPar;
    za.zPos = dZPos;
    za.zType = dZType;
    za.zCount = dZCount;
    za.status = dStatus;
    HIPCHK(launch_zero(za, stream));
  }
  mark(1);
  if (anyBim) {
    sb.gatherSrc = dPT[cur];
    sb.gatherDst = dSortedPT;
    sb.gatherDeadKey = deadKey;
    sb.sizeHint = (uint32_t)std::min<uint64_t>(knownN, 0xFFFFFFFFu);
    // the sort's last pass puts each particle's index into its sorted record
    // when the gather is fused into it and every index fits
    sb.packIndex = (sb.algo == 1 && cap <= kMaxPackedIndex) ? 1 : 0;
    HIPCHK(sort_keys(maxN, tilesLaid));
    mark(2);
    HIPCHK(launch_first(sb.keys, dN, maxN, nb, dFirst, dLive, dTables, slotsMode ? slotBits : 0u, R, stream));
    mark(3);
    CandArgs ca;
    ca.tables = dTables;
    ca.bucketStart = dBucketStart;
    ca.bucketRep = dBucketRep;
    ca.repShift = repShift;
    ca.R = (int)R;
    ca.keys = sb.keys;
    ca.sortedPT = dSortedPT;
    ca.first = dFirst;
    ca.dLive = dLive;
    ca.cand = dCand;
    HIPCHK(launch_candidates(ca, maxN, pairBlock, stream));
    mark(4);
    ResolveArgs ra;
    ra.tables = dTables;
    ra.streamReplica = dStreamReplica;
    ra.mtState = dMtState;
    ra.mtPar = dMtPar;
    ra.Stot = Stot;
    ra.first = dFirst;
    ra.cand = dCand;
    ra.sortedPT = dSortedPT;
    ra.vals = sb.vals;
    ra.packed = sb.packIndex;
    ra.pt = dPT[cur];
    ra.w = dW[cur];
    ra.repDeaths = dRepDeaths;
    ra.tileAlive = dTileAlive;
    HIPCHK(launch_resolve(ra, stream));
    mark(5);
  } else {
    for (int s = 2; s <= 5; ++s) mark(s);
  }
  FinalizeArgs fa;
  fa.R = (int)R;
  fa.repCount = dRepCount;
  fa.repDeaths = dRepDeaths;
  fa.repBirths = dRepBirths;
  fa.newOff = dNewOff;
  fa.newCount = dNewCount;
  fa.aliveBase = dAliveBase;
  fa.birthBase = dBirthBase;
  fa.repAlive = dRepAlive;
  fa.dNnext = dNnext;
  fa.dN = dN;
  fa.tileAlive = dTileAlive;
  fa.tileBirth = dTileBirth;
  fa.tileAlivePre = dTileAlivePre;
  fa.tileBirthPre = dTileBirthPre;
  fa.particleSteps = dParticleSteps;
  fa.stepCount = dStepCount;
  fa.countRing = dCountRing;
  fa.ringMask = kRing - 1;
  fa.injCount = dInjCount;
  fa.zCount = anyZero ? dZCount : nullptr;
  fa.zeroBase = dZeroBase;
  fa.injBase = dInjBase;
  fa.retire = dRetire;
  fa.blockSum = dBlockSum;
  fa.blockStride = blockStride;
  HIPCHK(launch_finalize(fa, maxN, stream));
  mark(6);
  ScatterArgs sa;
  sa.R = (int)R;
  sa.dN = dN;
  sa.repOff = dRepOff;
  sa.newOff = dNewOff;
  sa.aliveBase = dAliveBase;
  sa.birthBase = dBirthBase;
  sa.repAlive = dRepAlive;
  sa.tileAlivePre = dTileAlivePre;
  sa.tileBirthPre = dTileBirthPre;
  sa.pt = dPT[cur];
  sa.w = dW[cur];
  sa.addedPos = dAddedPos;
  sa.birthType = dBirthType;
  sa.tileBirth = dTileBirth;
  sa.ptOut = dPT[cur ^ 1];
  sa.wOut = dW[cur ^ 1];
  sa.blockSum = dBlockSum;
  sa.blockStride = blockStride;
  HIPCHK(launch_scatter(sa, maxN, stream));
  if (anyZero)
    HIPCHK(launch_zero_place((int)R, dTables, dZPos, dZType, dZCount, dNewOff, dZeroBase, dPT[cur ^ 1],
                             dW[cur ^ 1], stream));
  if (injN) {
    InjectArgs ia;
    ia.injPos = dInjPos;
    ia.injType = dInjType;
    ia.injRep = dInjRep;
    ia.injSrcOff = dInjSrcOff;
    ia.nInj = injN;
    ia.newOff = dNewOff;
    ia.injBase = dInjBase;
    ia.ptOut = dPT[cur ^ 1];
    ia.wOut = dW[cur ^ 1];
    HIPCHK(launch_inject(ia, stream));
    injN = 0;
  }
  // every admission (empty ones too) has entered with this step
  std::fill(slotPending.begin(), slotPending.end(), (uint8_t)0);
  mark(7);
  cur ^= 1;
  std::swap(dRepOff, dNewOff);
  std::swap(dRepCount, dNewCount);
  std::swap(dN, dNnext);
  return SMD_OK;
}

smd_status Engine::step(int nsteps) {
  if (!finalized) {
    set_error("engine not finalized");
    return SMD_ERR_INVALID;
  }
  if (broken) {
    set_error("engine state was consumed by a debug stage call");
    return SMD_ERR_INVALID;
  }
  HIPCHK(hipSetDevice(device));
  if (injN && nsteps > 0) {
    // admitted particles enter through an eager step (graphs hold no injection)
    smd_status st = ensure_capacity();
    if (st != SMD_OK) return st;
    st = enqueue_step(nullptr);
    if (st != SMD_OK) return st;
    --nsteps;
  }
  if (useGraphs && nsteps >= 2) {
    smd_status st = step_graph(nsteps / 2);
    if (st != SMD_OK) return st;
    nsteps -= 2 * (nsteps / 2);
  }
  for (int s = 0; s < nsteps; ++s) {
    smd_status st = ensure_capacity();
    if (st != SMD_OK) return st;
    st = enqueue_step(nullptr);
    if (st != SMD_OK) return st;
  }
  return SMD_OK;
}

smd_status Engine::sync() {
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipStreamSynchronize(stream));
  refresh_bound(true);
  return check_status();
}

// Error bits the kernels set since finalize (they stay set: the state after
// such a step is not trustworthy).
smd_status Engine::check_status() {
  if (!dStatus) return SMD_OK;
  uint32_t v = 0;
  HIPCHK(hipMemcpy(&v, dStatus, sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (v & kStatusZeroOverflow) {
    set_error("order-0 reactions created more particles in one step than the per-step limit");
    return SMD_ERR_CAPACITY;
  }
  if (v & kStatusSortSpin) {
    set_error("sort look-back wait exceeded its bound");
    return SMD_ERR_HIP;
  }
  if (v & kStatusSortLayout) {
    set_error("segmented sort: a destination fell outside the particle array");
    return SMD_ERR_HIP;
  }
  return SMD_OK;
}

smd_status Engine::set_knob(int knob, int value) {
  int* slot = nullptr;
  bool ok = false;
  if (knob == SMD_KNOB_PRODUCER_BLOCK) {
    slot = &producerBlock;
    ok = value == 128 || value == 256 || value == 512;
  } else if (knob == SMD_KNOB_PAIR_BLOCK) {
    slot = &pairBlock;
    ok = value == 64 || value == 128 || value == 256;
  } else if (knob == SMD_KNOB_SEGMENTED_SORT) {
    slot = &segSort;
    ok = value == 0 || value == 1;
  }
  if (!slot || !ok) {
    set_error("unknown knob or value out of range");
    return SMD_ERR_INVALID;
  }
  if (*slot != value) {
    drop_graph();  // captured launches keep the old shape
    *slot = value;
  }
  return SMD_OK;
}

// Serving layout: nSlots replicas with fixed bucket ranges (2^slotBucketBits
// each) and stream ranges (slotStreams each), all empty at first.
smd_status Engine::finalize_slots(uint32_t nSlots, uint32_t slotBucketBits, uint32_t slotStreamsMax,
                                  uint32_t capacity, uint32_t injectCap) {
  if (finalized || !tabs.empty()) {
    set_error("finalize_slots needs a fresh engine");
    return SMD_ERR_INVALID;
  }
  if (nSlots == 0 || slotBucketBits < 6 || slotBucketBits > 24 || slotStreamsMax == 0 ||
      ((uint64_t)nSlots << slotBucketBits) >= (1ull << 31)) {
    set_error("bad slot layout");
    return SMD_ERR_INVALID;
  }
  // Placeholder replicas: no particles, no streams, no reactions.
  smd_sim_spec s;
  std::memset(&s, 0, sizeof(s));
  s.num_types = 2;
  s.dt = 1.0f;
  for (int d = 0; d < 3; ++d) {
    s.box_hi[d] = 1.0f;
    s.grid_bits[d] = 2;
  }
  for (int f = 0; f < 6; ++f) s.face[f] = kReflect;
  s.num_streams = slotStreamsMax;
  for (uint32_t r = 0; r < nSlots; ++r) {
    const int id = add_replica(s, nullptr, nullptr, 0);
    if (id < 0) return (smd_status)(-id);
    tabs[r].numBuckets = 1u << slotBucketBits;  // reserve the whole range
  }
  slotsMode = true;
  slotBits = slotBucketBits;
  slotStreams = slotStreamsMax;
  injCap = injectCap;
  smd_status st = finalize(capacity);
  if (st != SMD_OK) return st;
  anySplit = true;  // jobs may create particles
  anyBim = true;
  // placeholders own no streams
  std::vector<int> none(Stot, -1);
  HIPCHK(hipMemcpy(dStreamReplica, none.data(), Stot * sizeof(int), hipMemcpyHostToDevice));
  // placeholders own no streams and no buckets (their ranges stay reserved)
  for (uint32_t r = 0; r < R; ++r) {
    tabs[r].numStreams = 0;
    tabs[r].numBuckets = 0;
  }
  HIPCHK(hipMemcpy(dTables, tabs.data(), R * sizeof(ReplicaTables), hipMemcpyHostToDevice));
  for (uint32_t** p : {&dRetire, &dInjCount, &dInjBase, &dInjSrcOff}) {
    HIPCHK(dalloc(p, R));
    HIPCHK(hipMemset(*p, 0, R * sizeof(uint32_t)));
  }
  HIPCHK(dalloc(&dInjPos, injCap));
  HIPCHK(dalloc(&dInjType, injCap));
  HIPCHK(dalloc(&dInjRep, injCap));
  slotBusy.assign(R, 0);
  slotPending.assign(R, 0);
  return SMD_OK;
}

smd_status Engine::admit(int slot, const smd_sim_spec& spec, const float* pos4, const int* types, uint32_t n) {
  if (!slotsMode || slot < 0 || (uint32_t)slot >= R) {
    set_error("admit needs a slot engine and a valid slot");
    return SMD_ERR_INVALID;
  }
  if (slotBusy[slot] || slotPending[slot]) {
    set_error("slot is occupied (retire its job first, and step once after an admission)");
    return SMD_ERR_INVALID;
  }
  if (injN + n > injCap) {
    set_error("admission exceeds the injection buffer of this step");
    return SMD_ERR_CAPACITY;
  }
  ReplicaTables T;
  std::string err;
  std::vector<smd_rxn_derived> der(spec.num_reactions > 0 ? spec.num_reactions : 1);
  if (!derive_tables(spec, T, der.data(), err)) {
    set_error(err);
    return SMD_ERR_RXN_SETUP;
  }
  if (T.numZero) {
    set_error("order-0 reactions are not supported in slot engines");
    return SMD_ERR_INVALID;
  }
  if (T.gbits[0] + T.gbits[1] + T.gbits[2] > slotBits || T.numStreams > slotStreams) {
    set_error("job grid or stream count exceeds the slot");
    return SMD_ERR_INVALID;
  }
  for (uint32_t i = 0; i < n; ++i)
    if (types[i] < 1 || types[i] >= spec.num_types) {
      set_error("particle type out of range");
      return SMD_ERR_INVALID;
    }
  smd_status st = sync();
  if (st != SMD_OK) return st;
  const uint32_t S = T.numStreams;
  T.bucketOff = tabs[slot].bucketOff;
  T.streamOff = tabs[slot].streamOff;
  tabs[slot] = T;
  specs[slot] = spec;
  derived[slot] = der;
  HIPCHK(hipMemcpy(dTables + slot, &T, sizeof(ReplicaTables), hipMemcpyHostToDevice));
  std::vector<int> rep(round64(slotStreams), -1);
  for (uint32_t k = 0; k < round64(S); ++k) rep[k] = slot;
  HIPCHK(hipMemcpy(dStreamReplica + T.streamOff, rep.data(), rep.size() * sizeof(int), hipMemcpyHostToDevice));
  const uint32_t* dat = mt_dat(err);
  if (!dat) {
    set_error(err);
    return SMD_ERR_IO;
  }
  std::vector<uint32_t> seeds(S), par(S);
  if (!stream_seeds(spec.seed, S, seeds.data(), err)) {
    set_error(err);
    return SMD_ERR_IO;
  }
  for (int w = 0; w < 3; ++w) {
    for (uint32_t k = 0; k < S; ++k) par[k] = dat[4 * (k % kMtParamSets) + w];
    HIPCHK(hipMemcpy(dMtPar + (size_t)w * Stot + T.streamOff, par.data(), S * sizeof(uint32_t),
                     hipMemcpyHostToDevice));
  }
  uint32_t* dSeeds = nullptr;
  HIPCHK(dalloc(&dSeeds, S));
  HIPCHK(hipMemcpy(dSeeds, seeds.data(), S * sizeof(uint32_t), hipMemcpyHostToDevice));
  HIPCHK(launch_mt_init_range(dMtState, dSeeds, Stot, T.streamOff, S, stream));
  HIPCHK(hipStreamSynchronize(stream));
  dfree(dSeeds);
  // particles enter during the next step's compaction
  if (n) {
    HIPCHK(hipMemcpy(dInjPos + injN, pos4, n * sizeof(float4), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dInjType + injN, types, n * sizeof(int), hipMemcpyHostToDevice));
    std::vector<int> r(n, slot);
    HIPCHK(hipMemcpy(dInjRep + injN, r.data(), n * sizeof(int), hipMemcpyHostToDevice));
  }
  HIPCHK(hipMemcpy(dInjSrcOff + slot, &injN, sizeof(uint32_t), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dInjCount + slot, &n, sizeof(uint32_t), hipMemcpyHostToDevice));
  injN += n;
  Nub += n;
  slotBusy[slot] = 1;
  slotPending[slot] = 1;
  return SMD_OK;
}

smd_status Engine::retire(int slot) {
  if (!slotsMode || slot < 0 || (uint32_t)slot >= R) {
    set_error("retire needs a slot engine and a valid slot");
    return SMD_ERR_INVALID;
  }
  if (slotPending[slot]) {
    set_error("the slot's job has not entered yet (step once after an admission)");
    return SMD_ERR_INVALID;
  }
  smd_status st = sync();
  if (st != SMD_OK) return st;
  const uint32_t one = 1u;
  HIPCHK(hipMemcpy(dRetire + slot, &one, sizeof(uint32_t), hipMemcpyHostToDevice));
  slotBusy[slot] = 0;
  return SMD_OK;
}

smd_status Engine::set_particles(const float* pos4, const int* types, uint32_t n) {
  if (!finalized || R != 1) {
    set_error("set_particles needs a finalized single-replica engine");
    return SMD_ERR_INVALID;
  }
  for (uint32_t i = 0; i < n; ++i)
    if (types[i] < 1 || types[i] >= specs[0].num_types) {
      set_error("particle type out of range");
      return SMD_ERR_INVALID;
    }
  smd_status st = sync();
  if (st != SMD_OK) return st;
  if (step_bound(n) > cap) {
    knownN = 0;
    st = grow(4ull * n);
    if (st != SMD_OK) return st;
  }
  if (n) HIPCHK(upload_particles(pos4, types, n, dPT[cur], dW[cur]));
  const uint32_t off[2] = {0u, n};
  HIPCHK(hipMemcpy(dRepOff, off, 2 * sizeof(uint32_t), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dRepCount, &n, sizeof(uint32_t), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dN, &n, sizeof(uint32_t), hipMemcpyHostToDevice));
  Nub = n;
  knownN = n;
  lastSeen = stepsDone;  // counts published before this call are stale
  return SMD_OK;
}

smd_status Engine::particle_steps(unsigned long long* out) {
  smd_status st = sync();
  if (st != SMD_OK) return st;
  HIPCHK(hipMemcpy(out, dParticleSteps, sizeof(unsigned long long), hipMemcpyDeviceToHost));
  return SMD_OK;
}

smd_status Engine::counts(uint32_t* out) {
  smd_status st = sync();
  if (st != SMD_OK) return st;
  HIPCHK(hipMemcpy(out, dRepCount, R * sizeof(uint32_t), hipMemcpyDeviceToHost));
  return SMD_OK;
}

smd_status Engine::get(int r, float* pos4, int* types, uint32_t capN, uint32_t* nOut) {
  if (r < 0 || (uint32_t)r >= R) {
    set_error("replica id out of range");
    return SMD_ERR_INVALID;
  }
  smd_status st = sync();
  if (st != SMD_OK) return st;
  uint32_t off = 0, n = 0;
  HIPCHK(hipMemcpy(&off, dRepOff + r, sizeof(uint32_t), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(&n, dRepCount + r, sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (nOut) *nOut = n;
  const uint32_t m = std::min(n, capN);
  if (m && (pos4 || types)) HIPCHK(download_particles(dPT[cur] + off, dW[cur] + off, m, pos4, types));
  return SMD_OK;
}

smd_status Engine::view(float4** pos4, int** types, uint32_t* nOut) {
  if (!finalized) {
    set_error("engine not finalized");
    return SMD_ERR_INVALID;
  }
  smd_status st = sync();
  if (st != SMD_OK) return st;
  uint32_t n = 0;
  HIPCHK(hipMemcpy(&n, dN, sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (n > viewCap || !dViewPos) {
    dfree(dViewPos);
    dfree(dViewType);
    viewCap = std::max<uint32_t>(n, 1u);
    HIPCHK(dalloc(&dViewPos, viewCap));
    HIPCHK(dalloc(&dViewType, viewCap));
  }
  HIPCHK(launch_unpack(dPT[cur], dW[cur], n, dViewPos, dViewType, stream));
  HIPCHK(hipStreamSynchronize(stream));
  *pos4 = dViewPos;
  *types = dViewType;
  *nOut = n;
  return SMD_OK;
}

smd_status Engine::get_streams(int r, uint32_t* words, uint32_t capStreams, bool set) {
  if (r < 0 || (uint32_t)r >= R) {
    set_error("replica id out of range");
    return SMD_ERR_INVALID;
  }
  smd_status st = sync();
  if (st != SMD_OK) return st;
  const ReplicaTables& T = tabs[r];
  const uint32_t m = std::min(T.numStreams, capStreams);
  std::vector<uint32_t> buf((size_t)m);
  for (int w = 0; w < kMtWords; ++w) {
    uint32_t* dsrc = dMtState + (size_t)w * Stot + T.streamOff;
    if (set) {
      for (uint32_t k = 0; k < m; ++k) buf[k] = words[(size_t)k * kMtWords + w];
      HIPCHK(hipMemcpy(dsrc, buf.data(), m * sizeof(uint32_t), hipMemcpyHostToDevice));
    } else {
      HIPCHK(hipMemcpy(buf.data(), dsrc, m * sizeof(uint32_t), hipMemcpyDeviceToHost));
      for (uint32_t k = 0; k < m; ++k) words[(size_t)k * kMtWords + w] = buf[k];
    }
  }
  return SMD_OK;
}

smd_status Engine::debug_stage(int stage, float* pos4, int* types, uint32_t* keysSorted, uint32_t* valsSorted,
                               int* cand, uint32_t* nTotal, uint32_t* nLive) {
  if (!finalized || broken) {
    set_error("engine not usable");
    return SMD_ERR_INVALID;
  }
  HIPCHK(hipSetDevice(device));
  smd_status st = sync();
  if (st != SMD_OK) return st;
  uint32_t n = 0;
  HIPCHK(hipMemcpy(&n, dN, sizeof(uint32_t), hipMemcpyDeviceToHost));
  const uint32_t maxN = n;
  broken = true;
  ProducerArgs pa;
  pa.tables = dTables;
  pa.streamReplica = dStreamReplica;
  pa.repOff = dRepOff;
  pa.repCount = dRepCount;
  pa.mtState = dMtState;
  pa.mtPar = dMtPar;
  pa.Stot = Stot;
  pa.gauss = dGauss;
  pa.pt = dPT[cur];
  pa.w = dW[cur];
  pa.addedPos = dAddedPos;
  pa.birthType = dBirthType;
  pa.key = sb.keys;
  pa.deadKey = deadKey;
  pa.repDeaths = dRepDeaths;
  pa.repBirths = dRepBirths;
  pa.tileAlive = dTileAlive;
  pa.tileBirth = dTileBirth;
  pa.retire = nullptr;
  HIPCHK(launch_producer(pa, producerBlock, stream));
  uint32_t live = 0;
  if (stage >= 2) {
    sb.gatherSrc = dPT[cur];
    sb.gatherDst = dSortedPT;
    sb.gatherDeadKey = deadKey;
    sb.sizeHint = (uint32_t)std::min<uint64_t>(knownN, 0xFFFFFFFFu);
    sb.packIndex = 0;  // the sorted values are returned
    HIPCHK(sort_keys(maxN));
    HIPCHK(launch_first(sb.keys, dN, maxN, nb, dFirst, dLive, dTables, slotsMode ? slotBits : 0u, R, stream));
  }
  if (stage >= 3) {
    CandArgs ca;
    ca.tables = dTables;
    ca.bucketStart = dBucketStart;
    ca.bucketRep = dBucketRep;
    ca.repShift = repShift;
    ca.R = (int)R;
    ca.keys = sb.keys;
    ca.sortedPT = dSortedPT;
    ca.first = dFirst;
    ca.dLive = dLive;
    ca.cand = dCand;
    HIPCHK(launch_candidates(ca, maxN, pairBlock, stream));
  }
  if (stage >= 4) {
    ResolveArgs ra;
    ra.tables = dTables;
    ra.streamReplica = dStreamReplica;
    ra.mtState = dMtState;
    ra.mtPar = dMtPar;
    ra.Stot = Stot;
    ra.first = dFirst;
    ra.cand = dCand;
    ra.sortedPT = dSortedPT;
    ra.vals = sb.vals;
    ra.packed = 0;
    ra.pt = dPT[cur];
    ra.w = dW[cur];
    ra.repDeaths = dRepDeaths;
    ra.tileAlive = dTileAlive;
    HIPCHK(launch_resolve(ra, stream));
  }
  HIPCHK(hipStreamSynchronize(stream));
  if (stage >= 2) HIPCHK(hipMemcpy(&live, dLive, sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (nTotal) *nTotal = n;
  if (nLive) *nLive = live;
  if (n) {
    if (pos4 || types) HIPCHK(download_particles(dPT[cur], dW[cur], n, pos4, types));
    if (stage == 1 && keysSorted) HIPCHK(hipMemcpy(keysSorted, sb.keys, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (stage >= 2 && keysSorted)
      HIPCHK(hipMemcpy(keysSorted, sb.keys, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (stage >= 2 && valsSorted)
      HIPCHK(hipMemcpy(valsSorted, sb.vals, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (stage >= 3 && cand && live) HIPCHK(hipMemcpy(cand, dCand, live * sizeof(int), hipMemcpyDeviceToHost));
  }
  return SMD_OK;
}

smd_status Engine::profile(int nsteps, double* us, int maxStages, char* names, int namesCap, int* nStages) {
  static const char* kNames[7] = {"producer", "sort", "first", "candidates", "resolve", "finalize", "scatter"};
  if (!finalized || broken) {
    set_error("engine not usable");
    return SMD_ERR_INVALID;
  }
  HIPCHK(hipSetDevice(device));
  StageTimer tm;
  for (int k = 0; k < 8; ++k) HIPCHK(hipEventCreate(&tm.ev[k]));
  std::vector<double> acc(7, 0.0);
  for (int s = 0; s < nsteps; ++s) {
    smd_status st = ensure_capacity();
    if (st != SMD_OK) return st;
    st = enqueue_step(&tm);
    if (st != SMD_OK) return st;
    HIPCHK(hipEventSynchronize(tm.ev[7]));
    for (int k = 0; k < 7; ++k) {
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, tm.ev[k], tm.ev[k + 1]));
      acc[k] += ms * 1000.0;
    }
  }
  for (int k = 0; k < 8; ++k) (void)hipEventDestroy(tm.ev[k]);
  const int m = std::min(7, maxStages);
  for (int k = 0; k < m; ++k) us[k] = nsteps > 0 ? acc[k] / nsteps : 0.0;
  if (nStages) *nStages = m;
  if (names && namesCap > 0) {
    std::string s;
    for (int k = 0; k < m; ++k) {
      s += kNames[k];
      if (k + 1 < m) s += "\n";
    }
    std::strncpy(names, s.c_str(), (size_t)namesCap - 1);
    names[namesCap - 1] = 0;
  }
  return SMD_OK;
}

}  // namespace smd
