// Drives the ParticleSystem C++ class (include/particle_system.h) directly,
// the way code written against the reference's particleSystem.h does: the
// reference's uint3/float3 types (HIP's here) in the constructor, getters and
// randPos, the public reactionOrderInfo / reactionList members, update() and
// getArray().  Test helper: tests/test_gpu_particle_system.py writes the
// input file and compares the output with the CPU oracle.
//
//   ps_direct <in> <out>
//   in:  smd_sim_spec bytes, u32 n, u32 steps, n float4 positions, n int types
//   out: u32 n, n float4 positions, n int types after `steps` update() calls,
//        then 3 x 8 randPos floats, then per order (0, 1, 2): reactionsPerOrder,
//        localToGlobalSize, numberOfReactionsPerType[num_types],
//        reactionsPerTypeIdx[num_types], localToGlobal[size],
//        reactionIndexes[reactionsPerOrder]; then per reaction: order, nprod,
//        reactants (2), products (2) as int and bindRadiusSquared, prob,
//        productPositions (6) as float.  Ints and floats are 4 bytes.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <vector>

#include "particle_system.h"

static bool read_all(FILE* f, void* p, size_t n) { return fread(p, 1, n, f) == n; }

int main(int argc, char** argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: ps_direct <in> <out>\n");
    return 2;
  }
  FILE* in = fopen(argv[1], "rb");
  if (!in) return 2;
  smd_sim_spec spec;
  uint32_t n = 0, steps = 0;
  if (!read_all(in, &spec, sizeof(spec)) || !read_all(in, &n, 4) || !read_all(in, &steps, 4)) return 2;
  std::vector<float> pos(4 * (size_t)n);
  std::vector<int> types(n);
  if (!read_all(in, pos.data(), 16 * (size_t)n) || !read_all(in, types.data(), 4 * (size_t)n)) return 2;
  fclose(in);

  // the reference constructor with HIP's uint3, and the vector getters into
  // HIP's float3 / uint3
  {
    ParticleSystem ref(1024u, make_uint3(64, 64, 64), false);
    if (ref.lastStatus() != 0) return 3;
    const uint3 g = ref.getGridSize();
    const float3 o = ref.getWorldOrigin();
    const float3 c = ref.getCellSize();
    ref.setColliderPos(make_float3(0.25f, -0.5f, 0.75f));
    const float3 cp = ref.getColliderPos();
    if (g.x != 64 || g.y != 64 || g.z != 64 || o.x != -1.0f || c.y != 1.0f / 32.0f || cp.z != 0.75f) return 4;
  }

  ParticleSystem ps(spec, pos.data(), types.data(), n, 0);
  if (ps.lastStatus() != 0) {
    fprintf(stderr, "construct: %s\n", smd_last_error());
    return 5;
  }
  float3 rp[8];
  for (float3& p : rp) ps.randPos(p);
  for (uint32_t s = 0; s < steps; ++s) {
    if (!ps.update() && ps.lastStatus() != 0) {
      fprintf(stderr, "update: %s\n", smd_last_error());
      return 6;
    }
  }
  const uint32_t m = (uint32_t)ps.getNumParticles();
  const float* P = ps.getArray(ParticleSystem::POSITION);
  const int* T = ps.getTypes();
  FILE* out = fopen(argv[2], "wb");
  if (!out) return 7;
  fwrite(&m, 4, 1, out);
  fwrite(P, 16, m, out);
  fwrite(T, 4, m, out);
  for (const float3& p : rp) {
    const float xyz[3] = {p.x, p.y, p.z};
    fwrite(xyz, 4, 3, out);
  }
  const int nt = spec.num_types;
  for (int o = 0; o < 3; ++o) {
    const ReactionOrderInfo& info = ps.reactionOrderInfo[o];
    fwrite(&info.reactionsPerOrder, 4, 1, out);
    fwrite(&info.localToGlobalSize, 4, 1, out);
    fwrite(info.numberOfReactionsPerType, 4, nt, out);
    fwrite(info.reactionsPerTypeIdx, 4, nt, out);
    fwrite(info.localToGlobal, 4, info.localToGlobalSize, out);
    fwrite(info.reactionIndexes, 4, info.reactionsPerOrder, out);
  }
  const Reaction* rl = ps.reactionList;
  for (int r = 0; r < spec.num_reactions; ++r) {
    const int iv[6] = {rl->order[r], rl->nprod[r], rl->reactants[r].x, rl->reactants[r].y, rl->products[r].x,
                       rl->products[r].y};
    fwrite(iv, 4, 6, out);
    const float fv[2] = {rl->bindRadiusSquared[r], rl->prob[r]};
    fwrite(fv, 4, 2, out);
    fwrite(&rl->productPositions[6 * r], 4, 6, out);
  }
  fclose(out);
  return 0;
}
