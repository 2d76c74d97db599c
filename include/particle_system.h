// ParticleSystem: the reference's host-side simulation object
// (particleSystem.h:50-218) on top of the gfx950 engine.
//
// Same public surface: construction from (numParticles, gridSize, useGL),
// update(), reset(), getArray()/setArray() and the getters/setters.  The
// reference constructor builds its active hard-coded scenario (test3: two
// species, D = 0.3 and 0.1, dt = 0.01, no reactions; particleSystem.cpp:513-571)
// in the reference world [-1, 1]^3 with the clamp boundary; a second
// constructor takes any smd_sim_spec.  OpenGL is not supported
// (bUseOpenGL must be false; the VBO getters return 0).  The reference's
// uint3/float3 arguments and results are accepted as any x, y, z vector type.
#ifndef GPUSMOLDYN_PARTICLE_SYSTEM_H
#define GPUSMOLDYN_PARTICLE_SYSTEM_H

#include <cstdint>
#include <vector>

#include "gpusmoldyn.h"
#include "gpusmoldyn_legacy.h"

const int MAX_PARTICLES = 8388608;  // particleSystem.h:31

namespace smd {
struct Sfmt;
}

class ParticleSystem {
 public:
  enum ParticleConfig { CONFIG_RANDOM, CONFIG_GRID, _NUM_CONFIGS };
  enum ParticleArray { POSITION, VELOCITY };

  ParticleSystem(unsigned int numParticles, smd_u3 gridSize, bool bUseOpenGL);
  // the same with the reference's uint3 (any type with x, y, z members)
  template <typename U3, typename = decltype(((U3*)0)->x + ((U3*)0)->y + ((U3*)0)->z)>
  ParticleSystem(unsigned int numParticles, const U3& gridSize, bool bUseOpenGL)
      : ParticleSystem(numParticles, smd_u3{(unsigned)gridSize.x, (unsigned)gridSize.y, (unsigned)gridSize.z},
                       bUseOpenGL) {}
  ParticleSystem(const smd_sim_spec& spec, const float* pos4, const int* types, unsigned int n, int device = 0);
  ~ParticleSystem();
  ParticleSystem(const ParticleSystem&) = delete;
  ParticleSystem& operator=(const ParticleSystem&) = delete;

  // One timestep; false on extinction or more than MAX_PARTICLES particles
  // (particleSystem.cpp:961-969).
  bool update();
  void reset(ParticleConfig config);

  // Host mirror owned by the object, valid until the next call.  POSITION:
  // numParticles float4; VELOCITY: not simulated, returns nullptr.
  float* getArray(ParticleArray array);
  void setArray(ParticleArray array, const float* data, int start, int count);

  int getNumParticles() const { return (int)m_numParticles; }
  int getNumTypes() const { return m_spec.num_types; }
  float getDeltaTime() const { return m_spec.dt; }
  float getSystemVolume() const { return m_spec.volume; }
  unsigned int getPosBuffer() const { return 0; }
  unsigned int getTypesBuffer() const { return 0; }
  // device positions (float4) / types (int) of the current step: copies
  // that stay valid until the next update() or VBO call
  void* getCudaPosVBO() const;
  void* getCudaTypesVBO() const;

  void dumpGrid();
  void dumpParticles(unsigned int start, unsigned int count);

  void setCollideSpring(float x) { m_params.spring = x; }
  void setCollideDamping(float x) { m_params.damping = x; }
  void setCollideShear(float x) { m_params.shear = x; }
  void setCollideAttraction(float x) { m_params.attraction = x; }
  void setColliderPos(smd_f3 x) { m_params.colliderPos = x; }
  template <typename F3, typename = decltype(((F3*)0)->x + ((F3*)0)->y + ((F3*)0)->z)>
  void setColliderPos(const F3& x) {
    m_params.colliderPos = smd_f3{(float)x.x, (float)x.y, (float)x.z};
  }
  float getParticleRadius() { return m_params.particleRadius; }
  smd_f3 getColliderPos() { return m_params.colliderPos; }
  float getColliderRadius() { return m_params.colliderRadius; }
  smd_u3 getGridSize() { return m_params.gridSize; }
  smd_f3 getWorldOrigin() { return m_params.worldOrigin; }
  smd_f3 getCellSize() { return m_params.cellSize; }
  float* getDiffusionCoefficients() { return m_spec.D; }

  // A uniform point of the reference world (-1, 1)^3 (particleSystem.h:115-120):
  // unirandOOF(-1, 1) per coordinate from the object's SFMT-19937 generator,
  // seeded with 0 at construction as the reference's reaction setup does
  // (init_gen_rand(0), particleSystem.cpp:342, 437, 536).
  void randPos(smd_f3& pos);
  template <typename F3, typename = decltype(((F3*)0)->x + ((F3*)0)->y + ((F3*)0)->z)>
  void randPos(F3& pos) {
    smd_f3 p;
    randPos(p);
    pos.x = p.x;
    pos.y = p.y;
    pos.z = p.z;
  }

  // The reaction superstructure and reaction list (particleSystem.h:172-174,
  // reaction.h:36-66), host arrays owned by the object: per order the
  // reactions listed by reactant type (fillReactionOrderInfo,
  // smolreact.cpp:173-271), and per reaction the inputs and the parameters
  // rxnsetrate/rxnsetproduct derived (binding radius squared, unbinding
  // radius, probability, product offsets).  Read-only: the simulation uses
  // the specification it was built from.
  ReactionOrderInfo reactionOrderInfo[3];
  Reaction* reactionList;

  // Additions.
  // Grid, origin, cell size and particle radius of the hash/clamp (the
  // state and the random streams are kept).
  int setParams(const SimParams& p);
  const SimParams& getParams() const { return m_params; }
  const int* getTypes();
  std::vector<unsigned int> getSpeciesCounts();
  const smd_rxn_derived* getReactions() const { return m_rxn.data(); }
  int lastStatus() const { return m_status; }

 private:
  int rebuild(const float* pos4, const int* types, unsigned int n, const uint32_t* streams);
  void initGrid(unsigned int* size, float spacing, float jitter, unsigned int numParticles);
  void fillReactions();

  smd_engine* m_engine = nullptr;
  smd_sim_spec m_spec;
  SimParams m_params;
  int m_device = 0;
  unsigned int m_numParticles = 0;
  std::vector<float> m_hPos;
  std::vector<int> m_hTypes;
  std::vector<smd_rxn_derived> m_rxn;
  int m_status = 0;
  // storage behind reactionOrderInfo / reactionList
  Reaction m_reactionList;
  std::vector<int> m_rOrder, m_rNprod, m_oCount[3], m_oStart[3], m_oList[3], m_oIndexes[3];
  std::vector<float> m_rBind2, m_rUnbind, m_rPgem, m_rProb, m_rRate, m_rProdPos;
  std::vector<smd_i2> m_rProducts, m_rReactants;
  smd::Sfmt* m_rng = nullptr;  // randPos generator
};


#endif
