// Device-side building blocks shared by every kernel: the dcmt Mersenne
// Twister kept in LDS, the float helpers that turn its output into the
// reference's random variates, the rotation used to place reaction products,
// boundary handling and the uniform-grid hash.
//
// Every floating-point expression whose rounding matters is written with
// explicit fmaf() and the library is compiled with -ffp-contract=off, so the
// CPU oracle (compiled the same way) reproduces each bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "smd_types.h"

namespace smd {

// ---------------------------------------------------------------------------
// dcmt MT (MEXP 607): MersenneTwister.cuh:51-77.  The 19 state words of the
// lanes of a workgroup live in LDS, word-major: word w of lane t is at
// lds[w * BS + t], so a wave's accesses to one word index are bank-conflict
// free.  `mti1` caches lds[iState] exactly like the reference's state->mti1.
// ---------------------------------------------------------------------------
struct Mt {
  uint32_t* s;     // &lds[threadIdx.x]
  int i;           // iState
  uint32_t mti1;
  uint32_t a, mb, mc;
};

__device__ __forceinline__ int wrap19(int x) { return x >= kMtNN ? x - kMtNN : x; }

__device__ __forceinline__ uint32_t mt_temper(uint32_t x, uint32_t mb, uint32_t mc) {
  x ^= x >> 12;
  x ^= (x << 7) & mb;
  x ^= (x << 15) & mc;
  x ^= x >> 18;
  return x;
}

__device__ __forceinline__ uint32_t mt_twist(uint32_t cur, uint32_t next, uint32_t mid, uint32_t a) {
  uint32_t x = (cur & 0xFFFFFFFEu) | (next & 1u);
  return mid ^ (x >> 1) ^ ((x & 1u) ? a : 0u);
}

template <int BS>
__device__ __forceinline__ uint32_t mt_draw(Mt& m) {
  int i1 = wrap19(m.i + 1);
  int iM = wrap19(m.i + kMtMM);
  uint32_t next = m.s[i1 * BS];
  uint32_t mid = m.s[iM * BS];
  uint32_t x = mt_twist(m.mti1, next, mid, m.a);
  m.s[m.i * BS] = x;
  m.mti1 = next;
  m.i = i1;
  return mt_temper(x, m.mb, m.mc);
}

// Three consecutive draws with all six LDS reads issued up front.  Valid
// because draws k=1..3 read words i+k and i+k+9 (mod 19) and write words
// i..i+2; none of the reads hits a word written earlier in the batch.
template <int BS>
__device__ __forceinline__ void mt_draw3(Mt& m, uint32_t& o0, uint32_t& o1, uint32_t& o2) {
  const int i0 = m.i;
  const int i1 = wrap19(i0 + 1), i2 = wrap19(i0 + 2), i3 = wrap19(i0 + 3);
  const int m1 = wrap19(i0 + 9), m2 = wrap19(i0 + 10), m3 = wrap19(i0 + 11);
  const uint32_t r1 = m.s[i1 * BS], r2 = m.s[i2 * BS], r3 = m.s[i3 * BS];
  const uint32_t q1 = m.s[m1 * BS], q2 = m.s[m2 * BS], q3 = m.s[m3 * BS];
  const uint32_t x1 = mt_twist(m.mti1, r1, q1, m.a);
  const uint32_t x2 = mt_twist(r1, r2, q2, m.a);
  const uint32_t x3 = mt_twist(r2, r3, q3, m.a);
  m.s[i0 * BS] = x1;
  m.s[i1 * BS] = x2;
  m.s[i2 * BS] = x3;
  m.mti1 = r3;
  m.i = i3;
  o0 = mt_temper(x1, m.mb, m.mc);
  o1 = mt_temper(x2, m.mb, m.mc);
  o2 = mt_temper(x3, m.mb, m.mc);
}

// Global SoA stream state: word w of stream g at st[w * S + g] (w = 19 holds
// iState); per-stream parameters at par[k * S + g] (matrix_a, mask_b, mask_c).
template <int BS>
__device__ __forceinline__ void mt_load(Mt& m, uint32_t* lds_lane, const uint32_t* __restrict__ st,
                                        const uint32_t* __restrict__ par, uint32_t S, uint32_t g) {
  m.s = lds_lane;
#pragma unroll
  for (int w = 0; w < kMtNN; ++w) lds_lane[w * BS] = st[(size_t)w * S + g];
  m.i = (int)st[(size_t)kMtNN * S + g];
  m.mti1 = lds_lane[m.i * BS];
  m.a = par[g];
  m.mb = par[(size_t)S + g];
  m.mc = par[(size_t)2 * S + g];
}

template <int BS>
__device__ __forceinline__ void mt_store(const Mt& m, uint32_t* __restrict__ st, uint32_t S, uint32_t g) {
#pragma unroll
  for (int w = 0; w < kMtNN; ++w) st[(size_t)w * S + g] = m.s[w * BS];
  st[(size_t)kMtNN * S + g] = (uint32_t)m.i;
}

// ---------------------------------------------------------------------------
// Random variates (rand_kernel.cuh:32-51).
// ---------------------------------------------------------------------------
__device__ __forceinline__ float rand_real(uint32_t v) { return (float)v * (1.0f / 4294967296.0f); }

// sin and cos of x in [0, 2*pi] (Cephes sinf/cosf method: reduction by pi/4
// in three parts, minimax polynomials on [-pi/4, pi/4]).  The reference uses
// the hardware __sincosf (reaction_kernel.cuh:181-183); this is a fixed,
// portable evaluation so the GPU and CPU agree bit for bit.
__device__ __forceinline__ void smd_sincos(float x, float* s, float* c) {
  int j = (int)(x * 1.27323954473516f);
  j += (j & 1);
  const float y = (float)j;
  float r = fmaf(-y, 0.78515625f, x);
  r = fmaf(-y, 2.4187564849853515625e-4f, r);
  r = fmaf(-y, 3.77489497744594108e-8f, r);
  const float z = r * r;
  const float sp = fmaf(fmaf(fmaf(-1.9515295891e-4f, z, 8.3321608736e-3f), z, -1.6666654611e-1f) * z, r, r);
  const float cp = fmaf(fmaf(fmaf(2.443315711809948e-5f, z, -1.388731625493765e-3f), z, 4.166664568298827e-2f) * z, z,
                        fmaf(-0.5f, z, 1.0f));
  const int q = (j >> 1) & 3;
  float sv, cv;
  if (q == 0) { sv = sp; cv = cp; }
  else if (q == 1) { sv = cp; cv = -sp; }
  else if (q == 2) { sv = -sp; cv = -cp; }
  else { sv = -cp; cv = sp; }
  *s = sv;
  *c = cv;
}

// thetarandCCF = acos(1 - 2u) (rand_kernel.cuh:38-41), evaluated as
// 2 asin(sqrt(u)) for u <= 1/2 and pi - 2 asin(sqrt(1 - u)) above (1 - u is
// exact there), asin by the Cephes asinf polynomial.
__device__ __forceinline__ float smd_theta(float u) {
  const bool mirror = u > 0.5f;
  const float s = sqrtf(mirror ? 1.0f - u : u);
  float x, z;
  const bool big = s > 0.5f;
  if (big) {
    z = 0.5f * (1.0f - s);
    x = sqrtf(z);
  } else {
    x = s;
    z = x * x;
  }
  const float p = fmaf(fmaf(fmaf(fmaf(4.2163199048e-2f, z, 2.4181311049e-2f), z, 4.5470025998e-2f), z,
                            7.4953002686e-2f), z, 1.6666752422e-1f);
  float y = fmaf(p * z, x, x);
  if (big) y = 1.5707963267948966f - (y + y);
  const float th = y + y;
  return mirror ? 3.14159265358979f - th : th;
}

// DirCosM (reaction_kernel.cuh:177-200) followed by dotMV (:166-175) for the
// two product offsets of reaction pp[0..5].
__device__ __forceinline__ void rotate_offsets(float theta, float phi, float chi, const float* pp, float v1[3],
                                               float v2[3]) {
  float sp, cp, st, ct, sc, cc;
  smd_sincos(phi, &sp, &cp);
  smd_sincos(theta, &st, &ct);
  smd_sincos(chi, &sc, &cc);
  float m[9];
  m[0] = fmaf(cp * ct, cc, -(sp * sc));
  m[1] = fmaf(sp * ct, cc, cp * sc);
  m[2] = -st * cc;
  m[3] = fmaf(-cp * ct, sc, -(sp * cc));
  m[4] = fmaf(-sp * ct, sc, cp * cc);
  m[5] = st * sc;
  m[6] = cp * st;
  m[7] = sp * st;
  m[8] = ct;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    v1[i] = fmaf(m[3 * i + 2], pp[2], fmaf(m[3 * i + 1], pp[1], m[3 * i] * pp[0]));
    v2[i] = fmaf(m[3 * i + 2], pp[5], fmaf(m[3 * i + 1], pp[4], m[3 * i] * pp[3]));
  }
}

__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
  return fmaf(az, bz, fmaf(ay, by, ax * bx));
}

// Minimum-image component along a periodic axis.
// The image of p2 nearest p1: p2 itself unless the axis is periodic and the
// pair is more than half a box apart (then p2 shifted by one box length), so
// that non-wrapping pairs use p2 exactly as bimReact does.
__device__ __forceinline__ float nearest_image(float p1, float p2, const ReplicaTables& T, int axis) {
  if (T.periodic[axis]) {
    const float d = p1 - p2;
    if (d > T.halfL[axis]) return p2 + T.L[axis];
    if (d < -T.halfL[axis]) return p2 - T.L[axis];
  }
  return p2;
}

__device__ __forceinline__ float min_image(float d, const ReplicaTables& T, int axis) {
  if (T.periodic[axis]) {
    if (d > T.halfL[axis]) d -= T.L[axis];
    else if (d < -T.halfL[axis]) d += T.L[axis];
  }
  return d;
}

// ---------------------------------------------------------------------------
// Box faces (per axis: upper face first, then lower face, as in
// particles_kernel.cu:73-86) and reflective spheres.  Returns false when an
// absorbing face removes the particle.  Geo is ReplicaTables or BoxGeom.
// ---------------------------------------------------------------------------

// The box and grid fields of a replica's tables, copied once into registers
// by a loop that moves many particles of one replica (read through the
// scalar cache otherwise, once per particle).
struct BoxGeom {
  float lo[3], hi[3], L[3];
  int32_t face[6];
  float particleRadius;
  float origin[3], cellSize[3];
  uint32_t gbits[3];
};

__device__ __forceinline__ BoxGeom box_geom(const ReplicaTables& T) {
  BoxGeom G;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    G.lo[d] = T.lo[d];
    G.hi[d] = T.hi[d];
    G.L[d] = T.L[d];
    G.origin[d] = T.origin[d];
    G.cellSize[d] = T.cellSize[d];
    G.gbits[d] = T.gbits[d];
  }
#pragma unroll
  for (int f = 0; f < 6; ++f) G.face[f] = T.face[f];
  G.particleRadius = T.particleRadius;
  return G;
}

template <class Geo>
__device__ __forceinline__ bool apply_box(const Geo& T, float* p) {
  bool alive = true;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    float x = p[d];
    const float lo = T.lo[d], hi = T.hi[d];
    const int fl = T.face[2 * d], fh = T.face[2 * d + 1];
    if (fh == kPeriodic) {
      if (x >= hi) x -= T.L[d];
      else if (x < lo) x += T.L[d];
    } else {
      if (fh == kClamp) {
        if (x > hi - T.particleRadius) x = hi - T.particleRadius;
      } else if (x > hi) {
        if (fh == kReflect) x = hi + (hi - x);
        else alive = false;
      }
      if (fl == kClamp) {
        if (x < lo + T.particleRadius) x = lo + T.particleRadius;
      } else if (x < lo) {
        if (fl == kReflect) x = lo + (lo - x);
        else alive = false;
      }
    }
    p[d] = x;
  }
  return alive;
}

// Reflective sphere (semantics of surfaces_kernel.cuh:97-248: sidedness by
// |p-c| - R, crossing point on the old->new segment, specular reflection of
// the remaining displacement across the tangent plane at the crossing).
__device__ __forceinline__ void apply_spheres(const ReplicaTables& T, const float* o, float* p) {
  for (int k = 0; k < T.numSpheres; ++k) {
    const float cx = T.sphere[k][0], cy = T.sphere[k][1], cz = T.sphere[k][2], R = T.sphere[k][3];
    const float R2 = R * R;
    const float fx = o[0] - cx, fy = o[1] - cy, fz = o[2] - cz;
    const float gx = p[0] - cx, gy = p[1] - cy, gz = p[2] - cz;
    const float so = dot3(fx, fy, fz, fx, fy, fz) - R2;
    const float sn = dot3(gx, gy, gz, gx, gy, gz) - R2;
    const bool inOld = so < 0.0f, inNew = sn < 0.0f;
    if (inOld == inNew) continue;
    const float dx = p[0] - o[0], dy = p[1] - o[1], dz = p[2] - o[2];
    const float A = dot3(dx, dy, dz, dx, dy, dz);
    const float B = dot3(fx, fy, fz, dx, dy, dz);
    float disc = fmaf(B, B, -(A * so));
    if (disc < 0.0f) disc = 0.0f;
    const float sq = sqrtf(disc);
    float t = inOld ? (sq - B) / A : (-B - sq) / A;
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    const float qx = fmaf(t, dx, o[0]), qy = fmaf(t, dy, o[1]), qz = fmaf(t, dz, o[2]);
    const float nx = (qx - cx) / R, ny = (qy - cy) / R, nz = (qz - cz) / R;
    const float dn = dot3(p[0] - qx, p[1] - qy, p[2] - qz, nx, ny, nz);
    float rx = fmaf(-2.0f * dn, nx, p[0]);
    float ry = fmaf(-2.0f * dn, ny, p[1]);
    float rz = fmaf(-2.0f * dn, nz, p[2]);
    const float hx = rx - cx, hy = ry - cy, hz = rz - cz;
    const bool inRef = (dot3(hx, hy, hz, hx, hy, hz) - R2) < 0.0f;
    if (inRef != inOld) {
      // Rounding left the point on the far side: put it just inside the
      // original side, on the radius through the crossing point.
      const float f = inOld ? (1.0f - 1.0e-6f) : (1.0f + 1.0e-6f);
      rx = fmaf(nx * R, f, cx);
      ry = fmaf(ny * R, f, cy);
      rz = fmaf(nz * R, f, cz);
    }
    p[0] = rx;
    p[1] = ry;
    p[2] = rz;
  }
}

// calcGridPos + calcGridHash (particles_kernel.cu:169-186), local to a
// replica's grid; the caller adds the replica's bucket offset.
template <class Geo>
__device__ __forceinline__ uint32_t grid_hash(const Geo& T, float x, float y, float z) {
  const int gx = (int)floorf((x - T.origin[0]) / T.cellSize[0]);
  const int gy = (int)floorf((y - T.origin[1]) / T.cellSize[1]);
  const int gz = (int)floorf((z - T.origin[2]) / T.cellSize[2]);
  const uint32_t mx = (1u << T.gbits[0]) - 1u, my = (1u << T.gbits[1]) - 1u, mz = (1u << T.gbits[2]) - 1u;
  return (((uint32_t)gz & mz) << (T.gbits[0] + T.gbits[1])) | (((uint32_t)gy & my) << T.gbits[0]) |
         ((uint32_t)gx & mx);
}

__device__ __forceinline__ uint64_t lanemask_lt() {
  const uint32_t lane = threadIdx.x & 63u;
  return (lane == 0) ? 0ull : ((~0ull) >> (64u - lane));
}

__device__ __forceinline__ int pack_index_type(uint32_t index, int type) {
  return (int)((index << kIndexShift) | (uint32_t)type);
}

__device__ __forceinline__ uint32_t unpack_index(float w) { return __float_as_uint(w) >> kIndexShift; }

// The Gaussian table t (kGaussDim entries, stats.cpp:202-232) is
// antisymmetric and then rotated right by one: t[u] = -t[kGaussDim + 1 - u]
// for 2 <= u < kGaussDim and t[0] = -t[1].  Its half h[j] = t[j + 1],
// j < kGaussDim / 2, holds all of it: t[u] = h[u - 1] for 1 <= u <=
// kGaussDim / 2, otherwise -h[(kGaussDim - u) mod kGaussDim / 2].
__device__ __forceinline__ float gauss_entry(const float* h, uint32_t u) {
  const uint32_t k = u - 1u;  // wraps for u = 0
  const bool lower = k < (uint32_t)(kGaussDim / 2);
  const float v = h[lower ? k : (((uint32_t)kGaussDim - u) & (uint32_t)(kGaussDim / 2 - 1))];
  return lower ? v : -v;
}

// Number of set bits of m at lanes below this one (v_mbcnt).
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// The valid lanes of the wave whose BITS-bit digit equals this lane's.  For
// each digit bit b the lanes that agree with ours are ballot(bit b) when our
// bit is set and its complement when it is clear, that is ~(ballot ^ s) with
// s our bit spread over 32 bits; acc & ~(ballot ^ s) is one three-input
// bitwise operation per mask half (truth table over (acc, ballot, s): true
// at 100 and 111, 0x90).
template <int BITS>
__device__ __forceinline__ uint64_t wave_digit_peers(uint32_t d, bool valid) {
  const uint64_t v = __ballot(valid);
  uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
#pragma unroll
  for (int b = 0; b < BITS; ++b) {
    const uint32_t s = (uint32_t)((int32_t)(d << (31 - b)) >> 31);
    const uint64_t bb = __ballot(s != 0u);
    lo = __builtin_amdgcn_bitop3_b32(lo, (uint32_t)bb, s, 0x90);
    hi = __builtin_amdgcn_bitop3_b32(hi, (uint32_t)(bb >> 32), s, 0x90);
  }
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

}  // namespace smd
