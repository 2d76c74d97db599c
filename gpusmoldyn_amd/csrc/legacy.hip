// The reference's extern "C" layer (include/gpusmoldyn_legacy.h) on gfx950.
//
// Stage-by-stage entry points with the reference's arguments: the caller owns
// every array, positions are float4 per particle, RNG states are the 84-byte
// MersenneTwisterState records, reaction parameters are reached through a
// device Reaction struct of device pointers.  Process-wide state (SimParams,
// Gaussian table, dcmt parameter table) lives in device buffers per process,
// as the reference's __constant__ data does.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gpusmoldyn_legacy.h"
#include "smd_device.h"
#include "smd_kernels.h"
#include "smd_setup.h"

namespace smd {
namespace {

constexpr int kLegacyStreams = MT_RNG_COUNT;
constexpr int kAosWords = 21;  // mt[19], iState, mti1

struct LegacyCtx {
  std::mutex mu;
  int device = -1;
  SimParams params;
  bool haveParams = false;
  SimParams* dParams = nullptr;
  float* dGauss = nullptr;
  uint32_t* dMtTable = nullptr;  // 32768 records: matrix_a, mask_b, mask_c, seed
  void* scratch = nullptr;
  size_t scratchBytes = 0;
};

LegacyCtx& ctx() {
  static LegacyCtx c;
  return c;
}

int fail(hipError_t e, const char* what) {
  set_error(std::string("HIP error '") + hipGetErrorString(e) + "' in " + what);
  return SMD_ERR_HIP;
}

#define LCHK(x)                         \
  do {                                  \
    hipError_t e_ = (x);                \
    if (e_ != hipSuccess) return fail(e_, #x); \
  } while (0)

int ensure_device_state() {
  LegacyCtx& c = ctx();
  int dev = 0;
  LCHK(hipGetDevice(&dev));
  if (c.device == dev && c.dParams) return SMD_OK;
  c.device = dev;
  LCHK(hipMalloc((void**)&c.dParams, sizeof(SimParams)));
  LCHK(hipMemset(c.dParams, 0, sizeof(SimParams)));
  LCHK(hipMalloc((void**)&c.dGauss, kGaussDim * sizeof(float)));
  std::vector<float> g(kGaussDim);
  gauss_table(g.data(), kGaussDim);
  LCHK(hipMemcpy(c.dGauss, g.data(), kGaussDim * sizeof(float), hipMemcpyHostToDevice));
  LCHK(hipMalloc((void**)&c.dMtTable, 4 * sizeof(uint32_t) * kLegacyStreams));
  LCHK(hipMemset(c.dMtTable, 0, 4 * sizeof(uint32_t) * kLegacyStreams));
  return SMD_OK;
}

int scratch(size_t bytes, void** out) {
  LegacyCtx& c = ctx();
  if (c.scratchBytes < bytes) {
    if (c.scratch) (void)hipFree(c.scratch);
    c.scratch = nullptr;
    c.scratchBytes = 0;
    LCHK(hipMalloc(&c.scratch, bytes));
    c.scratchBytes = bytes;
  }
  *out = c.scratch;
  return SMD_OK;
}

inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// ---------------------------------------------------------------------------
// AoS stream records: load the 19 words into the workgroup's LDS state.
// ---------------------------------------------------------------------------
template <int BS>
__device__ __forceinline__ void aos_load(Mt& m, uint32_t* lds_lane, const uint32_t* rec, const uint32_t* table,
                                         uint32_t s) {
  m.s = lds_lane;
#pragma unroll
  for (int w = 0; w < kMtNN; ++w) lds_lane[w * BS] = rec[w];
  m.i = (int)rec[19];
  m.mti1 = rec[20];
  m.a = table[4 * s];
  m.mb = table[4 * s + 1];
  m.mc = table[4 * s + 2];
}

template <int BS>
__device__ __forceinline__ void aos_store(const Mt& m, uint32_t* rec) {
#pragma unroll
  for (int w = 0; w < kMtNN; ++w) rec[w] = m.s[w * BS];
  rec[19] = (uint32_t)m.i;
  rec[20] = m.mti1;
}

constexpr int kLBS = 256;

// MersenneTwisterInitialise per record (MersenneTwister.cuh:40-49, :80-91).
__global__ __launch_bounds__(256) void k_legacy_mt_init(uint32_t* st, const uint32_t* table) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= (uint32_t)kLegacyStreams) return;
  uint32_t* rec = st + (size_t)kAosWords * s;
  uint32_t v = table[4 * s + 3];
  rec[0] = v;
  for (int i = 1; i < kMtNN; ++i) {
    v = 1812433253u * (v ^ (v >> 30)) + (uint32_t)i;
    rec[i] = v;
  }
  rec[19] = 0;
  rec[20] = rec[0];
}

struct IntegrateArgs {
  const Reaction* rl;
  const int* l2g;
  const int* idx;
  const int* nrxn;
  int nrx1;
  float4* pos;
  int* types;
  float4* addedPos;
  int* addedTypes;
  int* birth;
  int* death;
  const float* rates;
  uint32_t* st;
  const uint32_t* table;
  const float* gauss;
  uint32_t n;
};

// diffuse (particles_kernel.cu:125-166) then uniReact
// (reaction_kernel.cuh:263-342) per particle; thread s owns stream s and
// handles particles s, s + 32768, ... in order.
__global__ __launch_bounds__(kLBS) void k_legacy_integrate(IntegrateArgs a) {
  __shared__ uint32_t lds_mt[kMtNN * kLBS];
  const uint32_t s = blockIdx.x * kLBS + threadIdx.x;
  if (s >= (uint32_t)kLegacyStreams || s >= a.n) return;
  Mt m;
  uint32_t* rec = a.st + (size_t)kAosWords * s;
  aos_load<kLBS>(m, &lds_mt[threadIdx.x], rec, a.table, s);
  for (uint32_t i = s; i < a.n; i += kLegacyStreams) {
    float4 p = a.pos[i];
    const int type = a.types[i];
    uint32_t u0, u1, u2;
    mt_draw3<kLBS>(m, u0, u1, u2);
    const float rate = a.rates[type > 0 ? type : 0];
    p.x = fmaf(rate, a.gauss[u0 & (kGaussDim - 1)], p.x);
    p.y = fmaf(rate, a.gauss[u1 & (kGaussDim - 1)], p.y);
    p.z = fmaf(rate, a.gauss[u2 & (kGaussDim - 1)], p.z);
    a.pos[i] = p;
    if (a.nrx1 <= 0) continue;
    a.birth[i] = 0;
    a.death[i] = 1;
    if (type <= 0) {
      a.death[i] = 0;
      continue;
    }
    for (int j = 0; j < a.nrxn[type]; ++j) {
      const int rxn = a.l2g[a.idx[type] + j];
      if (rand_real(mt_draw<kLBS>(m)) < a.rl->prob[rxn]) {
        const smd_i2 pr = a.rl->products[rxn];
        if (pr.x <= 0) {
          a.death[i] = 0;
        } else {
          a.types[i] = pr.x;
          if (pr.y > 0) {
            a.addedTypes[i] = pr.y;
            a.birth[i] = 1;
            a.addedPos[i] = p;
          }
        }
        break;
      }
    }
  }
  aos_store<kLBS>(m, rec);
}

// adjustPositions (particles_kernel.cu:60-95): clamp to [-1 + r, 1 - r].
__global__ __launch_bounds__(256) void k_legacy_adjust(float4* pos, uint32_t n, float r) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float4 p = pos[i];
  if (p.x > 1.0f - r) p.x = 1.0f - r;
  if (p.x < -1.0f + r) p.x = -1.0f + r;
  if (p.y > 1.0f - r) p.y = 1.0f - r;
  if (p.y < -1.0f + r) p.y = -1.0f + r;
  if (p.z > 1.0f - r) p.z = 1.0f - r;
  if (p.z < -1.0f + r) p.z = -1.0f + r;
  pos[i] = p;
}

__device__ __forceinline__ int3 legacy_grid_pos(const SimParams& P, float x, float y, float z) {
  return make_int3((int)floorf((x - P.worldOrigin.x) / P.cellSize.x), (int)floorf((y - P.worldOrigin.y) / P.cellSize.y),
                   (int)floorf((z - P.worldOrigin.z) / P.cellSize.z));
}

__device__ __forceinline__ uint32_t legacy_grid_hash(const SimParams& P, int3 g) {
  const uint32_t x = (uint32_t)g.x & (P.gridSize.x - 1u);
  const uint32_t y = (uint32_t)g.y & (P.gridSize.y - 1u);
  const uint32_t z = (uint32_t)g.z & (P.gridSize.z - 1u);
  return (z * P.gridSize.y) * P.gridSize.x + y * P.gridSize.x + x;
}

// calcHashD (particles_kernel.cu:188-207)
__global__ __launch_bounds__(256) void k_legacy_hash(uint32_t* hash, uint32_t* index, const float4* pos, uint32_t n,
                                                     const SimParams* P) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float4 p = pos[i];
  hash[i] = legacy_grid_hash(*P, legacy_grid_pos(*P, p.x, p.y, p.z));
  index[i] = i;
}

__global__ __launch_bounds__(256) void k_legacy_fill(uint32_t* p, uint32_t v, uint32_t n) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

// reorderDataAndFindCellStartD (particles_kernel.cu:209-272)
__global__ __launch_bounds__(256) void k_legacy_reorder(uint32_t* cellStart, uint32_t* cellEnd, float4* sortedPos,
                                                        const uint32_t* hash, const uint32_t* index,
                                                        const float4* oldPos, uint32_t n) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint32_t h = hash[i];
  const uint32_t prev = i > 0 ? hash[i - 1] : 0u;
  if (i == 0 || h != prev) {
    cellStart[h] = i;
    if (i > 0) cellEnd[prev] = i;
  }
  if (i == n - 1) cellEnd[h] = i + 1;
  sortedPos[i] = oldPos[index[i]];
}

struct CollideArgs {
  float4* newPos;
  const float4* sortedPos;
  const uint32_t* index;
  int* types;
  int* birthDeath;
  const float* D;
  const int* table;
  const Reaction* rl;
  const uint32_t* cellStart;
  const uint32_t* cellEnd;
  uint32_t n;
  uint32_t numTypes;
  const SimParams* P;
  int* cand;
  uint32_t* st;
  const uint32_t* mtTable;
};

// collideD/collideCell (particles_kernel.cu:316-497): nearest reactive
// partner of each sorted particle among the 27 neighbouring cells.
__global__ __launch_bounds__(256) void k_legacy_candidates(CollideArgs a) {
  const uint32_t j = blockIdx.x * 256 + threadIdx.x;
  if (j >= a.n) return;
  const SimParams& P = *a.P;
  const uint32_t m1 = a.index[j];
  const int t1 = a.types[m1];
  int best = -1;
  if (t1 <= 0) {
    a.birthDeath[m1] = 0;  // particles_kernel.cu:348-352
    a.cand[j] = -1;
    return;
  }
  const float4 p = a.sortedPos[j];
  float bestD2 = 3.4e38f;
  const int3 g = legacy_grid_pos(P, p.x, p.y, p.z);
  for (int z = -1; z <= 1; ++z)
    for (int y = -1; y <= 1; ++y)
      for (int x = -1; x <= 1; ++x) {
        const uint32_t h = legacy_grid_hash(P, make_int3(g.x + x, g.y + y, g.z + z));
        const uint32_t s0 = a.cellStart[h];
        if (s0 == 0xFFFFFFFFu) continue;
        const uint32_t e0 = a.cellEnd[h];
        for (uint32_t k = s0; k < e0; ++k) {
          if (k == j) continue;
          const uint32_t m2 = a.index[k];
          const int t2 = a.types[m2];
          if (t2 == -1) {
            a.birthDeath[m2] = 0;  // particles_kernel.cu:374-378
            continue;
          }
          const int r = a.table[t1 * (int)a.numTypes + t2];
          if (r == -1) continue;
          const float4 q = a.sortedPos[k];
          const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
          const float d2 = dot3(dx, dy, dz, dx, dy, dz);
          if (d2 <= a.rl->bindRadiusSquared[r] && (d2 < bestD2 || (d2 == bestD2 && (int)k < best))) {
            best = (int)k;
            bestD2 = d2;
          }
        }
      }
  a.cand[j] = best;
}

// Mutual-nearest pairs react (bimReact, reaction_kernel.cuh:202-258); sorted
// particle j draws from stream j % 32768.
__global__ __launch_bounds__(kLBS) void k_legacy_resolve(CollideArgs a) {
  __shared__ uint32_t lds_mt[kMtNN * kLBS];
  const uint32_t s = blockIdx.x * kLBS + threadIdx.x;
  if (s >= (uint32_t)kLegacyStreams || s >= a.n) return;
  Mt m;
  bool loaded = false;
  uint32_t* rec = a.st + (size_t)kAosWords * s;
  for (uint32_t j = s; j < a.n; j += kLegacyStreams) {
    const int c = a.cand[j];
    if (c <= (int)j || a.cand[c] != (int)j) continue;
    if (!loaded) {
      aos_load<kLBS>(m, &lds_mt[threadIdx.x], rec, a.mtTable, s);
      loaded = true;
    }
    const uint32_t m1 = a.index[j], m2 = a.index[c];
    const int t1 = a.types[m1], t2 = a.types[m2];
    const int r = a.table[t1 * (int)a.numTypes + t2];
    const float pr = a.rl->prob[r];
    if (!(pr == 1.0f || rand_real(mt_draw<kLBS>(m)) < pr)) continue;
    const float4 p1 = a.sortedPos[j], p2 = a.sortedPos[c];
    const float dc1 = a.D[t1], dc2 = a.D[t2];
    const float x = (dc1 == 0.0f && dc2 == 0.0f) ? 0.5f : dc2 / (dc1 + dc2);
    const float omx = 1.0f - x;
    const float rp[3] = {fmaf(x, p1.x, omx * p2.x), fmaf(x, p1.y, omx * p2.y), fmaf(x, p1.z, omx * p2.z)};
    const float th = smd_theta(rand_real(mt_draw<kLBS>(m)));
    const float ph = rand_real(mt_draw<kLBS>(m)) * 6.28318548f;
    const float ch = rand_real(mt_draw<kLBS>(m)) * 6.28318548f;
    float v1[3], v2[3];
    rotate_offsets(th, ph, ch, a.rl->productPositions + 6 * r, v1, v2);
    a.birthDeath[m1] = 1;
    a.birthDeath[m2] = a.rl->nprod[r] - 1;
    const smd_i2 prd = a.rl->products[r];
    a.types[m1] = prd.x;
    a.types[m2] = prd.y;
    a.newPos[m1] = make_float4(rp[0] + v1[0], rp[1] + v1[1], rp[2] + v1[2], 0.0f);
    a.newPos[m2] = make_float4(rp[0] + v2[0], rp[1] + v2[1], rp[2] + v2[2], 0.0f);
  }
  if (loaded) aos_store<kLBS>(m, rec);
}

// resetNonInteracting (particles_kernel.cu:499-508)
__global__ __launch_bounds__(256) void k_legacy_reset(int* birthDeath, uint32_t n) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n && birthDeath[i] == -1) birthDeath[i] = 1;
}

// compactPosAndTypeD (particles_kernel.cu:97-123), out of place.
__global__ __launch_bounds__(256) void k_legacy_compact(const int* scan, const int* flags, const float4* pos,
                                                        const int* types, float4* posOut, int* typesOut, uint32_t n) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n || flags[i] == 0) return;
  const int d = scan[i];
  posOut[d] = pos[i];
  typesOut[d] = types[i];
}

// Exclusive scan: per-1024 block scans, a scan of the block sums, then adds.
__global__ __launch_bounds__(256) void k_scan_blocks(const int* in, int* out, int* sums, uint32_t n) {
  __shared__ int wsum[4];
  const uint32_t base = blockIdx.x * 1024u;
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  int v[4];
  int tsum = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t e = base + threadIdx.x * 4u + (uint32_t)k;
    v[k] = e < n ? in[e] : 0;
    tsum += v[k];
  }
  int x = tsum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  int pre = 0;
  for (uint32_t k = 0; k < w; ++k) pre += wsum[k];
  int run = pre + x - tsum;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t e = base + threadIdx.x * 4u + (uint32_t)k;
    if (e < n) out[e] = run;
    run += v[k];
  }
  if (threadIdx.x == 255) sums[blockIdx.x] = pre + x;
}

__global__ __launch_bounds__(1024) void k_scan_sums(int* sums, uint32_t nb) {
  __shared__ int wsum[16];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  for (uint32_t b = 0; b < nb; b += 1024) {
    const uint32_t e = b + threadIdx.x;
    const int v = e < nb ? sums[e] : 0;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(x, o, 64);
      if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int pre = 0;
    for (uint32_t k = 0; k < w; ++k) pre += wsum[k];
    const int c0 = carry;
    if (e < nb) sums[e] = c0 + pre + x - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry = c0 + pre + x;
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_scan_add(int* out, const int* sums, uint32_t n) {
  const uint32_t e = blockIdx.x * 1024u + threadIdx.x * 4u;
  const int add = sums[blockIdx.x];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (e + (uint32_t)k < n) out[e + k] += add;
}

}  // namespace
}  // namespace smd

using namespace smd;

extern "C" {

int cudaInit(int argc, char** argv) {
  int dev = 0;
  for (int k = 1; k < argc; ++k)
    if (argv[k] && std::strncmp(argv[k], "-device=", 8) == 0) dev = std::atoi(argv[k] + 8);
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    set_error("no HIP device visible");
    return SMD_ERR_NO_DEVICE;
  }
  if (dev < 0 || dev >= n) {
    set_error("device index out of range");
    return SMD_ERR_INVALID;
  }
  LCHK(hipSetDevice(dev));
  std::lock_guard<std::mutex> lk(ctx().mu);
  return ensure_device_state();
}

int cudaGLInit(int, char**) {
  set_error("OpenGL interop is not supported");
  return SMD_ERR_UNSUPPORTED;
}

int allocateArray(void** devPtr, size_t size) {
  LCHK(hipMalloc(devPtr, size));
  return SMD_OK;
}

int freeArray(void* devPtr) {
  LCHK(hipFree(devPtr));
  return SMD_OK;
}

int threadSync(void) {
  LCHK(hipDeviceSynchronize());
  return SMD_OK;
}

int copyArrayFromDevice(void* host, const void* device, unsigned int vbo, int size) {
  if (vbo) {
    set_error("OpenGL buffer objects are not supported");
    return SMD_ERR_UNSUPPORTED;
  }
  LCHK(hipMemcpy(host, device, (size_t)size, hipMemcpyDeviceToHost));
  return SMD_OK;
}

int copyArrayToDevice(void* device, const void* host, int offset, int size) {
  LCHK(hipMemcpy((char*)device + offset, host, (size_t)size, hipMemcpyHostToDevice));
  return SMD_OK;
}

int registerGLBufferObject(unsigned int) {
  set_error("OpenGL buffer objects are not supported");
  return SMD_ERR_UNSUPPORTED;
}
int unregisterGLBufferObject(unsigned int) {
  set_error("OpenGL buffer objects are not supported");
  return SMD_ERR_UNSUPPORTED;
}
void* mapGLBufferObject(unsigned int) {
  set_error("OpenGL buffer objects are not supported");
  return nullptr;
}
int unmapGLBufferObject(unsigned int) {
  set_error("OpenGL buffer objects are not supported");
  return SMD_ERR_UNSUPPORTED;
}

int setParameters(SimParams* hostParams) {
  std::lock_guard<std::mutex> lk(ctx().mu);
  int st = ensure_device_state();
  if (st) return st;
  for (unsigned g : {hostParams->gridSize.x, hostParams->gridSize.y, hostParams->gridSize.z})
    if (g == 0 || (g & (g - 1))) {
      set_error("gridSize must be a power of two per axis (particles_kernel.cu:182)");
      return SMD_ERR_INVALID;
    }
  ctx().params = *hostParams;
  ctx().haveParams = true;
  LCHK(hipMemcpy(ctx().dParams, hostParams, sizeof(SimParams), hipMemcpyHostToDevice));
  return SMD_OK;
}

int copyGaussianToConstant(float* table, unsigned int dim) {
  std::lock_guard<std::mutex> lk(ctx().mu);
  int st = ensure_device_state();
  if (st) return st;
  if (!table || dim == 0 || dim > (unsigned)kGaussDim) {
    set_error("gaussian table must have 1..4096 entries");
    return SMD_ERR_INVALID;
  }
  LCHK(hipMemcpy(ctx().dGauss, table, dim * sizeof(float), hipMemcpyHostToDevice));
  return SMD_OK;
}

int smdInitializeRandomNumbersSeeded(unsigned int seed) {
  std::lock_guard<std::mutex> lk(ctx().mu);
  int st = ensure_device_state();
  if (st) return st;
  std::string err;
  const uint32_t* dat = mt_dat(err);
  if (!dat) {
    set_error(err);
    return SMD_ERR_IO;
  }
  std::vector<uint32_t> tab(4 * (size_t)kLegacyStreams), seeds(kLegacyStreams);
  if (!stream_seeds(seed, kLegacyStreams, seeds.data(), err)) {
    set_error(err);
    return SMD_ERR_IO;
  }
  for (int k = 0; k < kLegacyStreams; ++k) {
    tab[4 * k] = dat[4 * k];
    tab[4 * k + 1] = dat[4 * k + 1];
    tab[4 * k + 2] = dat[4 * k + 2];
    tab[4 * k + 3] = seeds[k];
  }
  LCHK(hipMemcpy(ctx().dMtTable, tab.data(), tab.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  return SMD_OK;
}

void InitializeRandomNumbers(void) { (void)smdInitializeRandomNumbersSeeded(4357u); }

int initializeMT(MersenneTwisterState* stateArray) {
  int st;
  {
    std::lock_guard<std::mutex> lk(ctx().mu);
    st = ensure_device_state();
  }
  if (st) return st;
  hipLaunchKernelGGL(k_legacy_mt_init, dim3(kLegacyStreams / 256), dim3(256), 0, nullptr, (uint32_t*)stateArray,
                     ctx().dMtTable);
  LCHK(hipGetLastError());
  return SMD_OK;
}

int integrateSystem(Reaction* reactionList, int* localToGlobalReaction1, int* reactionsPerTypeIdx1,
                    int* numberOfReactionsPerType1, int numberOfReactions1, float* pos, int* types, float* addedPos,
                    int* addedTypes, int* birthArray, int* deathArray, float* diffusionRates,
                    MersenneTwisterState* rngStateArray, float deltaTime, unsigned int numParticles,
                    unsigned int numTypes) {
  (void)deltaTime;
  (void)numTypes;
  if (numParticles == 0) return SMD_OK;
  IntegrateArgs a;
  a.rl = reactionList;
  a.l2g = localToGlobalReaction1;
  a.idx = reactionsPerTypeIdx1;
  a.nrxn = numberOfReactionsPerType1;
  a.nrx1 = numberOfReactions1;
  a.pos = (float4*)pos;
  a.types = types;
  a.addedPos = (float4*)addedPos;
  a.addedTypes = addedTypes;
  a.birth = birthArray;
  a.death = deathArray;
  a.rates = diffusionRates;
  a.st = (uint32_t*)rngStateArray;
  a.table = ctx().dMtTable;
  a.gauss = ctx().dGauss;
  a.n = numParticles;
  const uint32_t threads = numParticles < (uint32_t)kLegacyStreams ? numParticles : kLegacyStreams;
  hipLaunchKernelGGL(k_legacy_integrate, dim3(cdiv(threads, kLBS)), dim3(kLBS), 0, nullptr, a);
  LCHK(hipGetLastError());
  return SMD_OK;
}

int adjustPositions(float* pos, int* types, unsigned int numParticles, float particleRadius) {
  (void)types;
  if (numParticles == 0) return SMD_OK;
  hipLaunchKernelGGL(k_legacy_adjust, dim3(cdiv(numParticles, 256)), dim3(256), 0, nullptr, (float4*)pos,
                     numParticles, particleRadius);
  LCHK(hipGetLastError());
  return SMD_OK;
}

int calcHash(unsigned int* gridParticleHash, unsigned int* gridParticleIndex, float* pos, int numParticles) {
  if (!ctx().haveParams) {
    set_error("setParameters has not been called");
    return SMD_ERR_INVALID;
  }
  if (numParticles <= 0) return SMD_OK;
  hipLaunchKernelGGL(k_legacy_hash, dim3(cdiv(numParticles, 256)), dim3(256), 0, nullptr, gridParticleHash,
                     gridParticleIndex, (const float4*)pos, (uint32_t)numParticles, ctx().dParams);
  LCHK(hipGetLastError());
  return SMD_OK;
}

int reorderDataAndFindCellStart(unsigned int* cellStart, unsigned int* cellEnd, float* sortedPos,
                                unsigned int* gridParticleHash, unsigned int* gridParticleIndex, float* oldPos,
                                unsigned int numParticles, unsigned int numCells) {
  if (numCells)
    hipLaunchKernelGGL(k_legacy_fill, dim3(cdiv(numCells, 256)), dim3(256), 0, nullptr, cellStart, 0xFFFFFFFFu,
                       numCells);
  if (numParticles)
    hipLaunchKernelGGL(k_legacy_reorder, dim3(cdiv(numParticles, 256)), dim3(256), 0, nullptr, cellStart, cellEnd,
                       (float4*)sortedPos, gridParticleHash, gridParticleIndex, (const float4*)oldPos, numParticles);
  LCHK(hipGetLastError());
  return SMD_OK;
}

int collide(float* newPos, float* sortedPos, unsigned int* gridParticleIndex, int* types, int* birthDeath,
            float* diffusionCoefficients, int* reactionTable, Reaction* reactionList, unsigned int* cellStart,
            unsigned int* cellEnd, unsigned int numParticles, unsigned int numCells, unsigned int numTypes,
            MersenneTwisterState* rngStateArray) {
  (void)numCells;
  if (!ctx().haveParams) {
    set_error("setParameters has not been called");
    return SMD_ERR_INVALID;
  }
  if (numParticles == 0) return SMD_OK;
  void* sc = nullptr;
  int st;
  {
    std::lock_guard<std::mutex> lk(ctx().mu);
    st = scratch(sizeof(int) * (size_t)numParticles, &sc);
  }
  if (st) return st;
  CollideArgs a;
  a.newPos = (float4*)newPos;
  a.sortedPos = (const float4*)sortedPos;
  a.index = gridParticleIndex;
  a.types = types;
  a.birthDeath = birthDeath;
  a.D = diffusionCoefficients;
  a.table = reactionTable;
  a.rl = reactionList;
  a.cellStart = cellStart;
  a.cellEnd = cellEnd;
  a.n = numParticles;
  a.numTypes = numTypes;
  a.P = ctx().dParams;
  a.cand = (int*)sc;
  a.st = (uint32_t*)rngStateArray;
  a.mtTable = ctx().dMtTable;
  hipLaunchKernelGGL(k_legacy_candidates, dim3(cdiv(numParticles, 256)), dim3(256), 0, nullptr, a);
  const uint32_t threads = numParticles < (uint32_t)kLegacyStreams ? numParticles : kLegacyStreams;
  hipLaunchKernelGGL(k_legacy_resolve, dim3(cdiv(threads, kLBS)), dim3(kLBS), 0, nullptr, a);
  hipLaunchKernelGGL(k_legacy_reset, dim3(cdiv(numParticles, 256)), dim3(256), 0, nullptr, birthDeath, numParticles);
  LCHK(hipGetLastError());
  return SMD_OK;
}

int compactPosAndType(int* dCompactedTemp, int* deathBirth, float* dPos, int* dTypes, int numParticles) {
  if (numParticles <= 0) return 0;
  const size_t n = (size_t)numParticles;
  void* sc = nullptr;
  int st;
  {
    std::lock_guard<std::mutex> lk(ctx().mu);
    st = scratch(n * (sizeof(float4) + sizeof(int)), &sc);
  }
  if (st) return -st;
  float4* posOut = (float4*)sc;
  int* typesOut = (int*)(posOut + n);
  hipLaunchKernelGGL(k_legacy_compact, dim3(cdiv(n, 256)), dim3(256), 0, nullptr, dCompactedTemp, deathBirth,
                     (const float4*)dPos, dTypes, posOut, typesOut, (uint32_t)n);
  if (hipGetLastError() != hipSuccess) return -SMD_ERR_HIP;
  int lastFlag = 0, lastScan = 0;
  if (hipMemcpy(&lastFlag, deathBirth + n - 1, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(&lastScan, dCompactedTemp + n - 1, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
    return -SMD_ERR_HIP;
  const int count = lastFlag + lastScan;  // particleSystemCuda.cu:298-306
  if (count > 0) {
    if (hipMemcpy(dPos, posOut, (size_t)count * sizeof(float4), hipMemcpyDeviceToDevice) != hipSuccess ||
        hipMemcpy(dTypes, typesOut, (size_t)count * sizeof(int), hipMemcpyDeviceToDevice) != hipSuccess)
      return -SMD_ERR_HIP;
  }
  return count;
}

int radixSort(unsigned int* keys, unsigned int* values, unsigned int* tempKeys, unsigned int* tempValues,
              unsigned int* counters, unsigned int* countersSum, unsigned int* blockOffsets, void* scanPlan,
              unsigned int numElements, unsigned int keyBits, bool flipBits) {
  (void)counters;
  (void)countersSum;
  (void)blockOffsets;
  (void)scanPlan;
  if (flipBits) {
    set_error("flipBits (float keys) is not supported");
    return SMD_ERR_INVALID;
  }
  if (numElements == 0) return SMD_OK;
  if (keyBits == 0 || keyBits > 32) {
    set_error("keyBits must be in 1..32");
    return SMD_ERR_INVALID;
  }
  const int rbits = 8;
  const size_t countWords = sort_counts_words(numElements, rbits);
  void* sc = nullptr;
  int st;
  {
    std::lock_guard<std::mutex> lk(ctx().mu);
    st = scratch((countWords + 1) * sizeof(uint32_t), &sc);
  }
  if (st) return st;
  uint32_t* dN = (uint32_t*)sc;
  SortBuffers b;
  b.keys = keys;
  b.vals = values;
  b.keysAlt = tempKeys;
  b.valsAlt = tempValues;
  b.counts = dN + 1;
  b.status = nullptr;
  b.algo = 1;
  b.gatherSrc = nullptr;
  b.gatherDst = nullptr;
  b.gatherDeadKey = 0;
  b.sizeHint = 0;
  b.packIndex = 0;
  LCHK(hipMemcpy(dN, &numElements, sizeof(uint32_t), hipMemcpyHostToDevice));
  LCHK(launch_sort(b, dN, numElements, keyBits, rbits, false, nullptr));
  if (b.keys != keys) {
    LCHK(hipMemcpyAsync(keys, b.keys, numElements * sizeof(uint32_t), hipMemcpyDeviceToDevice, nullptr));
    LCHK(hipMemcpyAsync(values, b.vals, numElements * sizeof(uint32_t), hipMemcpyDeviceToDevice, nullptr));
  }
  return SMD_OK;
}

int smdExclusiveScanInt(int* out, const int* in, unsigned int n) {
  if (n == 0) return SMD_OK;
  const uint32_t nb = cdiv(n, 1024);
  void* sc = nullptr;
  int st;
  {
    std::lock_guard<std::mutex> lk(ctx().mu);
    st = scratch((size_t)nb * sizeof(int), &sc);
  }
  if (st) return st;
  int* sums = (int*)sc;
  hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(256), 0, nullptr, in, out, sums, n);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, nullptr, sums, nb);
  hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(256), 0, nullptr, out, sums, n);
  LCHK(hipGetLastError());
  return SMD_OK;
}

}  // extern "C"
