// Headless driver: the reference's particles.cpp in benchmark / QA mode
// (runBenchmark, particles.cpp:213-242; command line, :639-702) without
// OpenGL, on the ParticleSystem class of this library.
//
//   smd_particles [-n=N] [-grid=G] [-i=ITERATIONS] [-benchmark]
//                 [-qatest | -noprompt] [-device=D] [-counts]
//                 [-out=particles.bin] [-ref=ref_particles.bin]
//                 [-counts-every=K -counts-out=counts.csv]
//                 [-traj-every=K -traj=traj.bin]
//
// -benchmark runs ITERATIONS updates (default 300) and prints the
// reference's timing line.  -qatest/-noprompt runs one update (or
// ITERATIONS when given), writes the positions as raw float4 to -out, and
// when the -ref file exists compares them element-wise: an element is wrong
// when it differs by more than 5.0 (MAX_EPSILON_ERROR, particles.cpp:57); the
// test fails when more than 30% (THRESHOLD, :58) are wrong.  -counts prints
// the species histogram after the run.
//
// Result files (written when asked for, also with -benchmark or -qatest):
//   counts.csv  header "step,species_1,...,species_T", then one row after
//               step 0 (the initial state) and after every K-th update.
//   traj.bin    frames of the particle state after step 0 and every K-th
//               update; each frame is int32 step, int32 n, then n records of
//               float32 x, y, z and int32 type (little-endian, 16 bytes).
//               The file starts with the 8 bytes "SMDTRJ01".
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "particle_system.h"

namespace {

bool flag(int argc, char** argv, const char* name) {
  for (int k = 1; k < argc; ++k) {
    const char* a = argv[k];
    while (*a == '-') ++a;
    if (std::strcmp(a, name) == 0) return true;
  }
  return false;
}

bool value(int argc, char** argv, const char* name, std::string& out) {
  const size_t len = std::strlen(name);
  for (int k = 1; k < argc; ++k) {
    const char* a = argv[k];
    while (*a == '-') ++a;
    if (std::strncmp(a, name, len) == 0 && a[len] == '=') {
      out = a + len + 1;
      return true;
    }
  }
  return false;
}

int int_value(int argc, char** argv, const char* name, int def) {
  std::string v;
  return value(argc, argv, name, v) ? std::atoi(v.c_str()) : def;
}

}  // namespace

int main(int argc, char** argv) {
  const unsigned gridDim = (unsigned)int_value(argc, argv, "grid", 64);  // GRID_SIZE, particles.cpp:60
  const unsigned numParticles = (unsigned)int_value(argc, argv, "n", 4096);  // NUM_PARTICLES, :61
  const bool qa = flag(argc, argv, "qatest") || flag(argc, argv, "noprompt");
  const bool benchmark = flag(argc, argv, "benchmark");
  int iterations = int_value(argc, argv, "i", qa ? 1 : 0);
  std::string outPath = "particles.bin", refPath = "ref_particles.bin";
  value(argc, argv, "out", outPath);
  value(argc, argv, "ref", refPath);

  if (cudaInit(argc, argv) != 0) {
    std::fprintf(stderr, "cudaInit failed: %s\n", smd_last_error());
    return 2;
  }
  std::printf("grid: %u x %u x %u = %u cells\n", gridDim, gridDim, gridDim, gridDim * gridDim * gridDim);
  std::printf("particles: %u\n", numParticles);

  const smd_u3 gs = {gridDim, gridDim, gridDim};
  ParticleSystem ps(numParticles, gs, false);
  if (ps.lastStatus() != 0) {
    std::fprintf(stderr, "ParticleSystem: %s\n", smd_last_error());
    return 2;
  }
  ps.reset(ParticleSystem::CONFIG_GRID);  // initParticleSystem, particles.cpp:163-166
  if (!benchmark && !qa) {
    std::printf("nothing to do: give -benchmark or -qatest (this driver has no OpenGL view)\n");
    return 0;
  }
  if (iterations <= 0) iterations = 300;

  const int countsEvery = int_value(argc, argv, "counts-every", 0);
  const int trajEvery = int_value(argc, argv, "traj-every", 0);
  std::string countsPath = "counts.csv", trajPath = "traj.bin";
  value(argc, argv, "counts-out", countsPath);
  value(argc, argv, "traj", trajPath);
  FILE* countsFile = nullptr;
  FILE* trajFile = nullptr;
  if (countsEvery > 0) {
    countsFile = std::fopen(countsPath.c_str(), "w");
    if (!countsFile) {
      std::fprintf(stderr, "cannot write %s\n", countsPath.c_str());
      return 2;
    }
    std::fprintf(countsFile, "step");
    for (int t = 1; t < ps.getNumTypes(); ++t) std::fprintf(countsFile, ",species_%d", t);
    std::fprintf(countsFile, "\n");
  }
  if (trajEvery > 0) {
    trajFile = std::fopen(trajPath.c_str(), "wb");
    if (!trajFile || std::fwrite("SMDTRJ01", 1, 8, trajFile) != 8) {
      std::fprintf(stderr, "cannot write %s\n", trajPath.c_str());
      return 2;
    }
  }
  auto record = [&](int step) {
    if (countsFile && step % countsEvery == 0) {
      const std::vector<unsigned> c = ps.getSpeciesCounts();
      std::fprintf(countsFile, "%d", step);
      for (size_t t = 1; t < c.size(); ++t) std::fprintf(countsFile, ",%u", c[t]);
      std::fprintf(countsFile, "\n");
    }
    if (trajFile && step % trajEvery == 0) {
      const int32_t n = ps.getNumParticles();
      const float* pos = ps.getArray(ParticleSystem::POSITION);
      const int* types = ps.getTypes();
      const int32_t head[2] = {step, n};
      std::fwrite(head, sizeof(int32_t), 2, trajFile);
      for (int32_t k = 0; k < n; ++k) {
        float rec[4] = {pos[4 * k], pos[4 * k + 1], pos[4 * k + 2], 0.0f};
        std::memcpy(&rec[3], &types[k], sizeof(int32_t));
        std::fwrite(rec, sizeof(float), 4, trajFile);
      }
    }
  };
  record(0);

  const auto t0 = std::chrono::steady_clock::now();
  int done = 0;
  for (; done < iterations; ++done) {
    if (!ps.update()) break;
    if (countsFile || trajFile) record(done + 1);
  }
  threadSync();
  if (countsFile) std::fclose(countsFile);
  if (trajFile) std::fclose(trajFile);
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  std::printf("%d particles, total time for %d iterations: %0.3f ms\n", ps.getNumParticles(), done, ms);
  if (ps.lastStatus() != 0) std::fprintf(stderr, "update stopped: %s\n", smd_last_error());

  if (flag(argc, argv, "counts")) {
    const std::vector<unsigned> c = ps.getSpeciesCounts();
    for (size_t t = 1; t < c.size(); ++t) std::printf("species %zu: %u\n", t, c[t]);
  }

  int errors = 0;
  if (qa) {
    const int n = ps.getNumParticles();
    const float* pos = ps.getArray(ParticleSystem::POSITION);
    FILE* f = std::fopen(outPath.c_str(), "wb");
    if (!f || (n > 0 && std::fwrite(pos, sizeof(float) * 4, (size_t)n, f) != (size_t)n)) {
      std::fprintf(stderr, "cannot write %s\n", outPath.c_str());
      if (f) std::fclose(f);
      return 2;
    }
    std::fclose(f);
    FILE* r = std::fopen(refPath.c_str(), "rb");
    if (r) {
      std::vector<float> ref((size_t)4 * n);
      const size_t got = n > 0 ? std::fread(ref.data(), sizeof(float), ref.size(), r) : 0;
      std::fclose(r);
      size_t wrong = got == ref.size() ? 0 : ref.size();
      for (size_t k = 0; k < got && got == ref.size(); ++k)
        if (!(std::fabs(ref[k] - pos[k]) <= 5.0f)) ++wrong;
      if (ref.size() && (double)wrong / (double)ref.size() > 0.30) ++errors;
      std::printf("TEST %s\n", errors ? "FAILED!" : "PASSED!");
    } else {
      std::printf("wrote %s (%d particles); no %s to compare with\n", outPath.c_str(), n, refPath.c_str());
    }
  }
  return errors ? 1 : 0;
}
