// CPU oracle of the gpusmoldyn timestep -- test infrastructure only (see
// oracle.h for the followed reference code and the race-resolution rules).
// Compiled with -ffp-contract=off: the fused multiply-adds below are the
// explicit ones of the method.
#include "oracle.h"

#include <cmath>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const int kNN = 19;  // state words (MersenneTwister.h:30)
const int kMM = 9;   // MersenneTwister.cuh:29

inline float uniform01(uint32_t v) { return (float)v * (1.0f / 4294967296.0f); }  // rand_kernel.cuh:32-36

inline float dot3f(float ax, float ay, float az, float bx, float by, float bz) {
  return std::fma(az, bz, std::fma(ay, by, ax * bx));
}

struct Stream {
  uint32_t* st;  // 20 words
  const uint32_t* par;
  uint32_t draw() { return orc_mt_draw(st, par); }
};

}  // namespace

extern "C" {

void orc_mt_init(uint32_t* st, uint32_t seed) {
  // MersenneTwisterInitialise (MersenneTwister.cuh:40-49)
  st[0] = seed;
  for (int i = 1; i < kNN; ++i) st[i] = 1812433253u * (st[i - 1] ^ (st[i - 1] >> 30)) + (uint32_t)i;
  st[kNN] = 0;
}

uint32_t orc_mt_draw(uint32_t* st, const uint32_t* par) {
  // MersenneTwisterGenerate (MersenneTwister.cuh:51-77).  mti1 of the
  // reference always equals st[iState], so it is not stored separately.
  const int i = (int)st[kNN];
  const int i1 = (i + 1) % kNN;
  const int im = (i + kMM) % kNN;
  uint32_t y = (st[i] & 0xFFFFFFFEu) | (st[i1] & 0x1u);
  y = st[im] ^ (y >> 1) ^ ((y & 1u) ? par[0] : 0u);
  st[i] = y;
  st[kNN] = (uint32_t)i1;
  y ^= (y >> 12);
  y ^= (y << 7) & par[1];
  y ^= (y << 15) & par[2];
  y ^= (y >> 18);
  return y;
}

// Sine and cosine on [0, 2 pi]: Cody-Waite reduction by pi/4 in three parts
// and the Cephes single-precision minimax polynomials.
void orc_sincos(float x, float* s, float* c) {
  int oct = (int)(x * 1.27323954473516f);
  oct += (oct & 1);
  const float k = (float)oct;
  float r = std::fma(-k, 0.78515625f, x);
  r = std::fma(-k, 2.4187564849853515625e-4f, r);
  r = std::fma(-k, 3.77489497744594108e-8f, r);
  const float r2 = r * r;
  const float ps = std::fma(std::fma(std::fma(-1.9515295891e-4f, r2, 8.3321608736e-3f), r2, -1.6666654611e-1f) * r2, r, r);
  const float pc = std::fma(std::fma(std::fma(2.443315711809948e-5f, r2, -1.388731625493765e-3f), r2,
                                     4.166664568298827e-2f) * r2,
                            r2, std::fma(-0.5f, r2, 1.0f));
  switch ((oct >> 1) & 3) {
    case 0: *s = ps; *c = pc; break;
    case 1: *s = pc; *c = -ps; break;
    case 2: *s = -ps; *c = -pc; break;
    default: *s = -pc; *c = ps; break;
  }
}

// acos(1 - 2u) = 2 asin(sqrt(u)) for u <= 1/2 and pi - 2 asin(sqrt(1 - u))
// above (1 - u is exact there), asin by the Cephes asinf polynomial.
float orc_theta(float u) {
  const bool mirror = u > 0.5f;
  const float s = std::sqrt(mirror ? 1.0f - u : u);
  float x, z;
  const bool upper = s > 0.5f;
  if (upper) {
    z = 0.5f * (1.0f - s);
    x = std::sqrt(z);
  } else {
    x = s;
    z = x * x;
  }
  const float poly = std::fma(std::fma(std::fma(std::fma(4.2163199048e-2f, z, 2.4181311049e-2f), z, 4.5470025998e-2f), z,
                                       7.4953002686e-2f), z, 1.6666752422e-1f);
  float a = std::fma(poly * z, x, x);
  if (upper) a = 1.5707963267948966f - (a + a);
  const float th = a + a;
  return mirror ? 3.14159265358979f - th : th;
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

}  // extern "C"

namespace {

// DirCosM (reaction_kernel.cuh:177-200) and dotMV (:166-175) for both
// product offsets.
void rotated_offsets(float theta, float phi, float chi, const float* pp, float v1[3], float v2[3]) {
  float sp, cp, st, ct, sc, cc;
  orc_sincos(phi, &sp, &cp);
  orc_sincos(theta, &st, &ct);
  orc_sincos(chi, &sc, &cc);
  const float m[9] = {std::fma(cp * ct, cc, -(sp * sc)), std::fma(sp * ct, cc, cp * sc), -st * cc,
                      std::fma(-cp * ct, sc, -(sp * cc)), std::fma(-sp * ct, sc, cp * cc), st * sc,
                      cp * st, sp * st, ct};
  for (int row = 0; row < 3; ++row) {
    v1[row] = std::fma(m[3 * row + 2], pp[2], std::fma(m[3 * row + 1], pp[1], m[3 * row] * pp[0]));
    v2[row] = std::fma(m[3 * row + 2], pp[5], std::fma(m[3 * row + 1], pp[4], m[3 * row] * pp[3]));
  }
}

inline bool periodic_axis(const orc_net& n, int d) { return n.face[2 * d] == 3; }

// The image of p2 nearest p1: p2 itself unless the axis is periodic and the
// pair is more than half a box apart (then p2 shifted by one box length), so
// that non-wrapping pairs use p2 exactly as bimReact does.
inline float nearest_image(const orc_net& n, int d, float p1, float p2) {
  if (!periodic_axis(n, d)) return p2;
  const float L = n.hi[d] - n.lo[d];
  const float h = 0.5f * L;
  const float v = p1 - p2;
  if (v > h) return p2 + L;
  if (v < -h) return p2 - L;
  return p2;
}

inline float wrap_delta(const orc_net& n, int d, float v) {
  if (!periodic_axis(n, d)) return v;
  const float L = n.hi[d] - n.lo[d];
  const float h = 0.5f * L;
  if (v > h) return v - L;
  if (v < -h) return v + L;
  return v;
}

// Box faces: per axis the upper face, then the lower one (the order of
// particles_kernel.cu:73-86).  false = absorbed.
bool box_faces(const orc_net& n, float* p) {
  bool keep = true;
  for (int d = 0; d < 3; ++d) {
    const float lo = n.lo[d], hi = n.hi[d];
    const int fl = n.face[2 * d], fh = n.face[2 * d + 1];
    float x = p[d];
    if (fh == 3) {
      const float L = hi - lo;
      if (x >= hi) x -= L;
      else if (x < lo) x += L;
    } else {
      if (fh == 0) {
        if (x > hi - n.particle_radius) x = hi - n.particle_radius;
      } else if (x > hi) {
        if (fh == 1) x = hi + (hi - x);
        else keep = false;
      }
      if (fl == 0) {
        if (x < lo + n.particle_radius) x = lo + n.particle_radius;
      } else if (x < lo) {
        if (fl == 1) x = lo + (lo - x);
        else keep = false;
      }
    }
    p[d] = x;
  }
  return keep;
}

// Reflective spheres: find where old->new crosses, reflect the rest of the
// step across the tangent plane there.
void sphere_reflect(const orc_net& n, const float* o, float* p) {
  for (int k = 0; k < n.num_spheres; ++k) {
    const float* S = n.sphere[k];
    const float R2 = S[3] * S[3];
    const float fx = o[0] - S[0], fy = o[1] - S[1], fz = o[2] - S[2];
    const float gx = p[0] - S[0], gy = p[1] - S[1], gz = p[2] - S[2];
    const float sideOld = dot3f(fx, fy, fz, fx, fy, fz) - R2;
    const float sideNew = dot3f(gx, gy, gz, gx, gy, gz) - R2;
    const bool inOld = sideOld < 0.0f;
    if (inOld == (sideNew < 0.0f)) continue;
    const float dx = p[0] - o[0], dy = p[1] - o[1], dz = p[2] - o[2];
    const float A = dot3f(dx, dy, dz, dx, dy, dz);
    const float B = dot3f(fx, fy, fz, dx, dy, dz);
    float disc = std::fma(B, B, -(A * sideOld));
    if (disc < 0.0f) disc = 0.0f;
    const float root = std::sqrt(disc);
    float t = inOld ? (root - B) / A : (-B - root) / A;
    t = std::fmin(std::fmax(t, 0.0f), 1.0f);
    const float cx = std::fma(t, dx, o[0]), cy = std::fma(t, dy, o[1]), cz = std::fma(t, dz, o[2]);
    const float nx = (cx - S[0]) / S[3], ny = (cy - S[1]) / S[3], nz = (cz - S[2]) / S[3];
    const float along = dot3f(p[0] - cx, p[1] - cy, p[2] - cz, nx, ny, nz);
    float qx = std::fma(-2.0f * along, nx, p[0]);
    float qy = std::fma(-2.0f * along, ny, p[1]);
    float qz = std::fma(-2.0f * along, nz, p[2]);
    const float hx = qx - S[0], hy = qy - S[1], hz = qz - S[2];
    if (((dot3f(hx, hy, hz, hx, hy, hz) - R2) < 0.0f) != inOld) {
      const float f = inOld ? (1.0f - 1.0e-6f) : (1.0f + 1.0e-6f);
      qx = std::fma(nx * S[3], f, S[0]);
      qy = std::fma(ny * S[3], f, S[1]);
      qz = std::fma(nz * S[3], f, S[2]);
    }
    p[0] = qx;
    p[1] = qy;
    p[2] = qz;
  }
}

// calcGridPos + calcGridHash (particles_kernel.cu:169-186)
uint32_t cell_hash(const orc_net& n, const float* p) {
  uint32_t g[3];
  for (int d = 0; d < 3; ++d) {
    const int c = (int)std::floor((p[d] - n.origin[d]) / n.cell[d]);
    g[d] = (uint32_t)c & ((1u << n.gbits[d]) - 1u);
  }
  return (g[2] << (n.gbits[0] + n.gbits[1])) | (g[1] << n.gbits[0]) | g[0];
}

struct StepScratch {
  std::vector<uint32_t> key, order, first;
  std::vector<int> birthType, cand;
  std::vector<float> addedPos, spt;
  unsigned live = 0;
};

// Per particle: diffuse (3 draws) -> unimolecular reactions -> faces/spheres
// -> hash.  Streams are independent, so they run in parallel.
void producer(orc_sim& s, const float* gauss, StepScratch& w, uint32_t deadKey) {
  const orc_net& n = s.net;
  const unsigned N = s.n, S = n.num_streams;
  w.key.assign(N, deadKey);
  w.birthType.assign(N, 0);
  w.addedPos.assign(4 * (size_t)N, 0.0f);
  const long nS = (long)(S < N ? S : N);
#pragma omp parallel for schedule(static)
  for (long sid = 0; sid < nS; ++sid) {
    Stream rng{s.mt + 20 * (size_t)sid, s.par + 3 * (size_t)sid};
    for (unsigned i = (unsigned)sid; i < N; i += S) {
      float* P = s.pos + 4 * (size_t)i;
      int t = s.type[i];
      const uint32_t r0 = rng.draw(), r1 = rng.draw(), r2 = rng.draw();
      const float step = n.rate[t > 0 ? t : 0];
      float q[3] = {std::fma(step, gauss[r0 & 4095u], P[0]), std::fma(step, gauss[r1 & 4095u], P[1]),
                    std::fma(step, gauss[r2 & 4095u], P[2])};
      float w4 = P[3];
      int bt = 0;
      float b[3] = {q[0], q[1], q[2]};
      float bw = P[3];
      if (t <= 0) {
        t = 0;
      } else if (n.has_uni) {
        for (int j = 0; j < n.uni_count[t]; ++j) {
          const int rx = n.uni_list[n.uni_start[t] + j];
          if (!(uniform01(rng.draw()) < n.prob[rx])) continue;
          if (n.prod_x[rx] <= 0) {
            t = 0;
          } else {
            t = n.prod_x[rx];
            if (n.prod_y[rx] > 0) {
              bt = n.prod_y[rx];
              if (n.unbind_offsets) {
                const float th = orc_theta(uniform01(rng.draw()));
                const float ph = uniform01(rng.draw()) * 6.28318548f;
                const float ch = uniform01(rng.draw()) * 6.28318548f;
                float v1[3], v2[3];
                rotated_offsets(th, ph, ch, n.prod_pos[rx], v1, v2);
                for (int d = 0; d < 3; ++d) {
                  b[d] = q[d] + v2[d];
                  q[d] = q[d] + v1[d];
                }
                w4 = 0.0f;
                bw = 0.0f;
              }
            }
          }
          break;
        }
      }
      uint32_t key = deadKey;
      if (t > 0) {
        const float old[3] = {P[0], P[1], P[2]};
        if (box_faces(n, q)) {
          if (n.num_spheres) sphere_reflect(n, old, q);
          key = cell_hash(n, q);
        } else {
          t = 0;
        }
      }
      P[0] = q[0];
      P[1] = q[1];
      P[2] = q[2];
      P[3] = w4;
      s.type[i] = t;
      w.key[i] = key;
      w.birthType[i] = bt;
      if (bt) {
        float* A = &w.addedPos[4 * (size_t)i];
        A[0] = b[0];
        A[1] = b[1];
        A[2] = b[2];
        A[3] = bw;
      }
    }
  }
}

// Stable counting sort by bucket (the result of a stable radix sort of
// (hash, index) pairs); first[h] = number of keys < h.
void sort_and_cells(const orc_sim& s, StepScratch& w, uint32_t deadKey) {
  const unsigned N = s.n;
  w.first.assign((size_t)deadKey + 2, 0u);
  for (unsigned i = 0; i < N; ++i) w.first[w.key[i] + 1]++;
  for (size_t h = 1; h < w.first.size(); ++h) w.first[h] += w.first[h - 1];
  std::vector<uint32_t> fill(w.first.begin(), w.first.end() - 1);
  w.order.assign(N, 0u);
  for (unsigned i = 0; i < N; ++i) w.order[fill[w.key[i]]++] = i;
  w.live = w.first[deadKey];
  w.spt.assign(4 * (size_t)w.live, 0.0f);
  for (unsigned j = 0; j < w.live; ++j) {
    const unsigned i = w.order[j];
    w.spt[4 * (size_t)j] = s.pos[4 * (size_t)i];
    w.spt[4 * (size_t)j + 1] = s.pos[4 * (size_t)i + 1];
    w.spt[4 * (size_t)j + 2] = s.pos[4 * (size_t)i + 2];
    std::memcpy(&w.spt[4 * (size_t)j + 3], &s.type[i], sizeof(int));
  }
}

inline int spt_type(const float* spt, size_t j) {
  int t;
  std::memcpy(&t, &spt[4 * j + 3], sizeof(int));
  return t;
}

inline void consider(const orc_net& n, const float* spt, unsigned j, unsigned k, int& best, float& bestD2) {
  if (k == j) return;
  const int tj = spt_type(spt, j), tk = spt_type(spt, k);
  if (tk <= 0 || tk >= n.num_types) return;
  const int rx = n.pair_rxn[tj * ORC_MAX_TYPES + tk];
  if (rx < 0) return;
  const float dx = wrap_delta(n, 0, spt[4 * (size_t)j] - spt[4 * (size_t)k]);
  const float dy = wrap_delta(n, 1, spt[4 * (size_t)j + 1] - spt[4 * (size_t)k + 1]);
  const float dz = wrap_delta(n, 2, spt[4 * (size_t)j + 2] - spt[4 * (size_t)k + 2]);
  const float d2 = dot3f(dx, dy, dz, dx, dy, dz);
  if (d2 <= n.bind_r2[rx] && (d2 < bestD2 || (d2 == bestD2 && (int)k < best))) {
    best = (int)k;
    bestD2 = d2;
  }
}

// collideD's 27-cell walk (particles_kernel.cu:443-497), recording the
// nearest reactive partner instead of reacting on the spot.
void candidates(const orc_sim& s, StepScratch& w) {
  const orc_net& n = s.net;
  w.cand.assign(w.live, -1);
  if (!n.has_bim) return;
  const unsigned bx = n.gbits[0], by = n.gbits[1];
  const uint32_t mx = (1u << bx) - 1u, my = (1u << by) - 1u, mz = (1u << n.gbits[2]) - 1u;
  const long L = (long)w.live;
#pragma omp parallel for schedule(dynamic, 1024)
  for (long jj = 0; jj < L; ++jj) {
    const unsigned j = (unsigned)jj;
    const uint32_t key = w.key[w.order[j]];
    const uint32_t gx = key & mx, gy = (key >> bx) & my, gz = key >> (bx + by);
    int best = -1;
    float bestD2 = 3.4e38f;
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const uint32_t h = (((gz + (uint32_t)dz) & mz) << (bx + by)) | (((gy + (uint32_t)dy) & my) << bx) |
                             ((gx + (uint32_t)dx) & mx);
          for (uint32_t k = w.first[h]; k < w.first[h + 1]; ++k) consider(n, w.spt.data(), j, k, best, bestD2);
        }
    w.cand[j] = best;
  }
}

// Mutual-nearest pairs react (bimReact, reaction_kernel.cuh:202-258).
void resolve(orc_sim& s, StepScratch& w) {
  const orc_net& n = s.net;
  if (!n.has_bim) return;
  const unsigned S = n.num_streams;
  for (unsigned j = 0; j < w.live; ++j) {
    const int c = w.cand[j];
    if (c <= (int)j || w.cand[c] != (int)j) continue;
    Stream rng{s.mt + 20 * (size_t)(j % S), s.par + 3 * (size_t)(j % S)};
    const float* P1 = &w.spt[4 * (size_t)j];
    const float* P2 = &w.spt[4 * (size_t)c];
    const int t1 = spt_type(w.spt.data(), j), t2 = spt_type(w.spt.data(), (size_t)c);
    const int rx = n.pair_rxn[t1 * ORC_MAX_TYPES + t2];
    const float pr = n.prob[rx];
    if (!(pr == 1.0f || uniform01(rng.draw()) < pr)) continue;
    const float d1 = n.D[t1], d2 = n.D[t2];
    const float x = (d1 == 0.0f && d2 == 0.0f) ? 0.5f : d2 / (d1 + d2);
    const float omx = 1.0f - x;
    float at[3];
    for (int d = 0; d < 3; ++d) {
      const float img = nearest_image(n, d, P1[d], P2[d]);
      at[d] = std::fma(x, P1[d], omx * img);
    }
    const float th = orc_theta(uniform01(rng.draw()));
    const float ph = uniform01(rng.draw()) * 6.28318548f;
    const float ch = uniform01(rng.draw()) * 6.28318548f;
    float v1[3], v2[3];
    rotated_offsets(th, ph, ch, n.prod_pos[rx], v1, v2);
    const unsigned m1 = w.order[j], m2 = w.order[c];
    float* A = s.pos + 4 * (size_t)m1;
    float* B = s.pos + 4 * (size_t)m2;
    for (int d = 0; d < 3; ++d) {
      float a = at[d] + v1[d], b = at[d] + v2[d];
      if (periodic_axis(n, d)) {
        const float Lx = n.hi[d] - n.lo[d];
        if (a >= n.hi[d]) a -= Lx;
        else if (a < n.lo[d]) a += Lx;
        if (b >= n.hi[d]) b -= Lx;
        else if (b < n.lo[d]) b += Lx;
      }
      A[d] = a;
      B[d] = b;
    }
    A[3] = 0.0f;
    B[3] = 0.0f;
    s.type[m1] = n.prod_x[rx];
    s.type[m2] = n.prod_y[rx];
  }
}

// Survivors in index order, then births in parent index order.
int compact(orc_sim& s, StepScratch& w) {
  const unsigned N = s.n;
  unsigned alive = 0, births = 0;
  for (unsigned i = 0; i < N; ++i) {
    alive += s.type[i] > 0;
    births += w.birthType[i] > 0;
  }
  if (alive + births > s.cap) return 5;
  std::vector<float> np(4 * (size_t)(alive + births));
  std::vector<int> nt(alive + births);
  unsigned o = 0;
  for (unsigned i = 0; i < N; ++i)
    if (s.type[i] > 0) {
      std::memcpy(&np[4 * (size_t)o], &s.pos[4 * (size_t)i], 16);
      nt[o++] = s.type[i];
    }
  for (unsigned i = 0; i < N; ++i)
    if (w.birthType[i] > 0) {
      std::memcpy(&np[4 * (size_t)o], &w.addedPos[4 * (size_t)i], 16);
      nt[o++] = w.birthType[i];
    }
  std::memcpy(s.pos, np.data(), np.size() * sizeof(float));
  std::memcpy(s.type, nt.data(), nt.size() * sizeof(int));
  s.n = o;
  return 0;
}

// Order-0 creation after the compaction: for each order-0 reaction in index
// order, a Poisson count (product of closed-interval uniforms compared with
// exp(-mean)), then per new particle x, y, z = lo + u (hi - lo) with u on the
// open interval, w = 1.
int create_zeroth(orc_sim& s) {
  const orc_net& n = s.net;
  if (n.num_zero == 0) return 0;
  Stream z{s.zmt, s.zpar};
  auto closed = [&z]() { return (float)((double)z.draw() * (1.0 / 4294967295.0)); };
  auto open = [&z]() { return (float)(((double)z.draw() + 0.5) * (1.0 / 4294967296.0)); };
  for (int q = 0; q < n.num_zero; ++q) {
    const int rx = n.zero_list[q];
    unsigned nmol = 0;
    for (int p = 0; p < n.zero_pieces[q]; ++p)
      for (float t = closed(); t > n.zero_g[q]; t *= closed()) ++nmol;
    for (unsigned j = 0; j < nmol; ++j) {
      if (s.n >= s.cap) return 5;
      float* p = &s.pos[4 * (size_t)s.n];
      for (int d = 0; d < 3; ++d) p[d] = open() * (n.hi[d] - n.lo[d]) + n.lo[d];
      p[3] = 1.0f;
      s.type[s.n++] = n.prod_x[rx];
    }
  }
  return 0;
}

uint32_t dead_key(const orc_net& n) { return 1u << (n.gbits[0] + n.gbits[1] + n.gbits[2]); }

}  // namespace

extern "C" {

int orc_step(orc_sim* s, const float* gauss, int nsteps) {
  StepScratch w;
  const uint32_t dk = dead_key(s->net);
  for (int k = 0; k < nsteps; ++k) {
    producer(*s, gauss, w, dk);
    if (s->net.has_bim) {
      sort_and_cells(*s, w, dk);
      candidates(*s, w);
      resolve(*s, w);
    }
    int rc = compact(*s, w);
    if (rc) return rc;
    rc = create_zeroth(*s);
    if (rc) return rc;
  }
  return 0;
}

int orc_step_stage(orc_sim* s, const float* gauss, int stage, uint32_t* keys, uint32_t* vals, int* cand,
                   unsigned* n_live) {
  StepScratch w;
  const uint32_t dk = dead_key(s->net);
  producer(*s, gauss, w, dk);
  if (stage == 1) {
    if (keys) std::memcpy(keys, w.key.data(), w.key.size() * 4);
    return 0;
  }
  sort_and_cells(*s, w, dk);
  if (n_live) *n_live = w.live;
  if (keys)
    for (unsigned j = 0; j < s->n; ++j) keys[j] = w.key[w.order[j]];
  if (vals) std::memcpy(vals, w.order.data(), w.order.size() * 4);
  if (stage >= 3) {
    candidates(*s, w);
    if (cand) std::memcpy(cand, w.cand.data(), w.cand.size() * sizeof(int));
  }
  if (stage >= 4) resolve(*s, w);
  return 0;
}

void orc_candidates_bruteforce(const orc_net* net, const float* spt, unsigned n_live, int* cand) {
  const long L = (long)n_live;
#pragma omp parallel for schedule(dynamic, 64)
  for (long jj = 0; jj < L; ++jj) {
    int best = -1;
    float bestD2 = 3.4e38f;
    for (unsigned k = 0; k < n_live; ++k) consider(*net, spt, (unsigned)jj, k, best, bestD2);
    cand[jj] = best;
  }
}

}  // extern "C"
