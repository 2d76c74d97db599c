// ParticleSystem (include/particle_system.h) over the engine, and its C
// wrappers (smd_ps_* in include/gpusmoldyn.h).
#include "../../include/particle_system.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "smd_engine.h"
#include "smd_setup.h"

namespace {

unsigned ilog2(unsigned v) {
  unsigned b = 0;
  while ((1u << b) < v) ++b;
  return b;
}

float frand() { return rand() / (float)RAND_MAX; }  // particleSystem.cpp:1082-1085

// The reference's active scenario (test3Types/test3Reactions,
// particleSystem.cpp:513-571) in its world (particleSystem.cpp:62-85).
smd_sim_spec reference_spec(smd_u3 gridSize) {
  smd_sim_spec s;
  std::memset(&s, 0, sizeof(s));
  s.num_types = 3;
  s.D[0] = 0.0f;
  s.D[1] = 0.3f;
  s.D[2] = 0.1f;
  s.dt = 0.01f;
  s.volume = 90.0f;
  s.num_reactions = 0;
  const float r = 1.0f / 64.0f;
  for (int d = 0; d < 3; ++d) {
    s.box_lo[d] = -1.0f;
    s.box_hi[d] = 1.0f;
    s.cell_size[d] = r * 2.0f;
    s.origin[d] = -1.0f;
  }
  s.grid_bits[0] = ilog2(gridSize.x);
  s.grid_bits[1] = ilog2(gridSize.y);
  s.grid_bits[2] = ilog2(gridSize.z);
  for (int f = 0; f < 6; ++f) s.face[f] = SMD_BC_CLAMP;
  s.particle_radius = r;
  s.num_streams = MT_RNG_COUNT;
  s.seed = 4357u;
  return s;
}

SimParams params_from_spec(const smd_sim_spec& s) {
  SimParams p;
  std::memset(&p, 0, sizeof(p));
  p.particleRadius = s.particle_radius;
  p.colliderPos = {-1.2f, -0.8f, 0.8f};
  p.colliderRadius = 0.2f;
  p.gridSize = {1u << s.grid_bits[0], 1u << s.grid_bits[1], 1u << s.grid_bits[2]};
  p.numCells = p.gridSize.x * p.gridSize.y * p.gridSize.z;
  p.worldOrigin = {s.origin[0], s.origin[1], s.origin[2]};
  p.cellSize = {s.cell_size[0], s.cell_size[1], s.cell_size[2]};
  p.spring = 0.5f;
  p.damping = 0.02f;
  p.shear = 0.1f;
  p.attraction = 0.0f;
  p.boundaryDamping = -0.5f;
  p.gravity = {0.0f, -0.0003f, 0.0f};
  p.globalDamping = 1.0f;
  return p;
}

}  // namespace

ParticleSystem::ParticleSystem(unsigned int numParticles, smd_u3 gridSize, bool bUseOpenGL) {
  if (bUseOpenGL) {
    smd::set_error("OpenGL interop is not supported");
    m_status = SMD_ERR_UNSUPPORTED;
  }
  m_spec = reference_spec(gridSize);
  m_params = params_from_spec(m_spec);
  // positions start at 0 until reset() (m_hPos is zeroed, particleSystem.cpp:266-267);
  // the first 3/4 of the particles are species 1, the rest species 2
  std::vector<float> pos(4 * (size_t)numParticles, 0.0f);
  std::vector<int> types(numParticles);
  const unsigned bound = 3 * (numParticles / 4);
  for (unsigned i = 0; i < numParticles; ++i) types[i] = i < bound ? 1 : 2;
  if (m_status == 0) m_status = rebuild(pos.data(), types.data(), numParticles, nullptr);
  fillReactions();
}

ParticleSystem::ParticleSystem(const smd_sim_spec& spec, const float* pos4, const int* types, unsigned int n,
                               int device)
    : m_spec(spec), m_device(device) {
  m_params = params_from_spec(m_spec);
  m_status = rebuild(pos4, types, n, nullptr);
  fillReactions();
}

ParticleSystem::~ParticleSystem() {
  smd_engine_destroy(m_engine);
  delete m_rng;
}

void ParticleSystem::randPos(smd_f3& pos) {
  pos.x = m_rng->uniform_open(-1.0f, 1.0f);
  pos.y = m_rng->uniform_open(-1.0f, 1.0f);
  pos.z = m_rng->uniform_open(-1.0f, 1.0f);
}

// reactionList / reactionOrderInfo from the specification and the derived
// parameters (reaction.h:36-66; fillReactionOrderInfo, smolreact.cpp:173-271:
// per order, reactions listed by reactant type in global index order, an
// order-2 reaction under both of its reactants).
void ParticleSystem::fillReactions() {
  if (!m_rng) m_rng = new smd::Sfmt(0u);
  const int R = m_spec.num_reactions, T = m_spec.num_types;
  m_rOrder.assign(R, 0);
  m_rNprod.assign(R, 0);
  m_rBind2.assign(R, 0.0f);
  m_rUnbind.assign(R, 0.0f);
  m_rPgem.assign(R, 0.0f);
  m_rProb.assign(R, 0.0f);
  m_rRate.assign(R, 0.0f);
  m_rProdPos.assign(6 * (size_t)R, 0.0f);
  m_rProducts.assign(R, smd_i2{-1, -1});
  m_rReactants.assign(R, smd_i2{-1, -1});
  for (int r = 0; r < R; ++r) {
    const smd_reaction& x = m_spec.rxn[r];
    m_rOrder[r] = x.order;
    m_rNprod[r] = x.nprod;
    m_rRate[r] = x.rate;
    m_rReactants[r] = smd_i2{x.order >= 1 ? x.reactants[0] : -1, x.order == 2 ? x.reactants[1] : -1};
    m_rProducts[r] = smd_i2{x.nprod >= 1 ? x.products[0] : -1, x.nprod >= 2 ? x.products[1] : -1};
    m_rPgem[r] = x.pgem > 0.0f ? x.pgem : 0.2f;
    if ((size_t)r < m_rxn.size()) {
      m_rBind2[r] = m_rxn[r].bind_radius2;
      m_rUnbind[r] = m_rxn[r].unbind_radius;
      m_rProb[r] = m_rxn[r].prob;
      for (int q = 0; q < 6; ++q) m_rProdPos[6 * (size_t)r + q] = m_rxn[r].product_pos[q];
    }
  }
  m_reactionList.order = m_rOrder.data();
  m_reactionList.bindRadiusSquared = m_rBind2.data();
  m_reactionList.unbindRadius = m_rUnbind.data();
  m_reactionList.geminateProbability = m_rPgem.data();
  m_reactionList.prob = m_rProb.data();
  m_reactionList.rate = m_rRate.data();
  m_reactionList.nprod = m_rNprod.data();
  m_reactionList.products = m_rProducts.data();
  m_reactionList.reactants = m_rReactants.data();
  m_reactionList.productPositions = m_rProdPos.data();
  reactionList = &m_reactionList;
  for (int o = 0; o < 3; ++o) {
    m_oCount[o].assign(T, 0);
    m_oStart[o].assign(T, 0);
    m_oIndexes[o].clear();
    for (int r = 0; r < R; ++r) {
      const smd_reaction& x = m_spec.rxn[r];
      if (x.order != o) continue;
      m_oIndexes[o].push_back(r);
      if (o >= 1) m_oCount[o][x.reactants[0]]++;
      if (o == 2) m_oCount[o][x.reactants[1]]++;
    }
    for (int t = 1; t < T; ++t) m_oStart[o][t] = m_oStart[o][t - 1] + m_oCount[o][t - 1];
    const int total = T ? m_oStart[o][T - 1] + m_oCount[o][T - 1] : 0;
    m_oList[o].assign(total, 0);
    std::vector<int> fill = m_oStart[o];
    for (int r = 0; r < R && o >= 1; ++r) {
      const smd_reaction& x = m_spec.rxn[r];
      if (x.order != o) continue;
      m_oList[o][fill[x.reactants[0]]++] = r;
      if (o == 2) m_oList[o][fill[x.reactants[1]]++] = r;
    }
    ReactionOrderInfo& info = reactionOrderInfo[o];
    info.order = o;
    info.reactionsPerOrder = (int)m_oIndexes[o].size();
    info.reactionsPerTypeIdx = m_oStart[o].data();
    info.localToGlobal = m_oList[o].data();
    info.localToGlobalSize = total;
    info.numberOfReactionsPerType = m_oCount[o].data();
    info.reactionIndexes = m_oIndexes[o].data();
  }
}

int ParticleSystem::rebuild(const float* pos4, const int* types, unsigned int n, const uint32_t* streams) {
  smd_engine* e = smd_engine_create(m_device, 0);
  if (!e) return SMD_ERR_NO_DEVICE;
  const int r = smd_engine_add_replica(e, &m_spec, pos4, types, n);
  if (r < 0) {
    smd_engine_destroy(e);
    return -r;
  }
  smd_status st = smd_engine_finalize(e, 0);
  if (st == SMD_OK && streams) st = smd_engine_set_streams(e, 0, streams, e->impl->tabs[0].numStreams);
  if (st != SMD_OK) {
    smd_engine_destroy(e);
    return st;
  }
  smd_engine_destroy(m_engine);
  m_engine = e;
  m_numParticles = n;
  m_rxn.assign(e->impl->derived[0].begin(), e->impl->derived[0].end());
  return SMD_OK;
}

bool ParticleSystem::update() {
  if (!m_engine) return false;
  m_status = smd_engine_step(m_engine, 1);
  if (m_status != SMD_OK) return false;
  unsigned n = 0;
  m_status = smd_engine_counts(m_engine, &n);
  if (m_status != SMD_OK) return false;
  m_numParticles = n;
  return n > 0 && n <= (unsigned)MAX_PARTICLES;
}

float* ParticleSystem::getArray(ParticleArray array) {
  if (array != POSITION || !m_engine) return nullptr;
  unsigned n = 0;
  m_status = smd_engine_get(m_engine, 0, nullptr, nullptr, 0, &n);
  m_hPos.resize(4 * (size_t)std::max(n, 1u));
  m_status = smd_engine_get(m_engine, 0, m_hPos.data(), nullptr, n, &n);
  m_numParticles = n;
  return m_hPos.data();
}

const int* ParticleSystem::getTypes() {
  if (!m_engine) return nullptr;
  unsigned n = 0;
  m_status = smd_engine_get(m_engine, 0, nullptr, nullptr, 0, &n);
  m_hTypes.resize(std::max(n, 1u));
  m_status = smd_engine_get(m_engine, 0, nullptr, m_hTypes.data(), n, &n);
  return m_hTypes.data();
}

void ParticleSystem::setArray(ParticleArray array, const float* data, int start, int count) {
  if (array != POSITION || !m_engine || start < 0 || count < 0) return;
  unsigned n = 0;
  smd_engine_get(m_engine, 0, nullptr, nullptr, 0, &n);
  std::vector<float> pos(4 * (size_t)std::max(n, 1u));
  std::vector<int> types(std::max(n, 1u));
  smd_engine_get(m_engine, 0, pos.data(), types.data(), n, &n);
  const unsigned end = std::min<unsigned>(n, (unsigned)(start + count));
  for (unsigned i = (unsigned)start; i < end; ++i)
    for (int k = 0; k < 4; ++k) pos[4 * (size_t)i + k] = data[4 * (size_t)(i - start) + k];
  m_status = m_engine->impl->set_particles(pos.data(), types.data(), n);
}

void ParticleSystem::initGrid(unsigned int* size, float spacing, float jitter, unsigned int numParticles) {
  // particleSystem.cpp:1087-1116
  (void)spacing;
  srand(1973);
  const float zsp = 2.0f / (size[2] - 2), ysp = 2.0f / (size[1] - 2), xsp = 2.0f / (size[0] - 2);
  const float r = m_params.particleRadius;
  for (unsigned z = 0; z < size[2]; z++)
    for (unsigned y = 0; y < size[1]; y++)
      for (unsigned x = 0; x < size[0]; x++) {
        const unsigned i = (z * size[1] * size[0]) + (y * size[0]) + x;
        if (i < numParticles) {
          m_hPos[i * 4] = (xsp * x) + r - 1.0f + (frand() * 2.0f - 1.0f) * jitter;
          m_hPos[i * 4 + 1] = (ysp * y) + r - 1.0f + (frand() * 2.0f - 1.0f) * jitter;
          m_hPos[i * 4 + 2] = (zsp * z) + r - 1.0f + (frand() * 2.0f - 1.0f) * jitter;
          m_hPos[i * 4 + 3] = 1.0f;
        }
      }
}

void ParticleSystem::reset(ParticleConfig config) {
  // particleSystem.cpp:1118-1161; unlike the reference, positions are
  // uploaded whether or not OpenGL is used
  if (!m_engine) return;
  const unsigned n = m_numParticles;
  m_hPos.assign(4 * (size_t)std::max(n, 1u), 0.0f);
  if (config == CONFIG_GRID) {
    const float jitter = m_params.particleRadius * 0.01f;
    unsigned s = (unsigned)std::ceil(std::pow((float)n, 1.0f / 3.0f));
    unsigned gs[3] = {s, s, s};
    initGrid(gs, m_params.particleRadius * 2.0f, jitter, n);
  } else {
    for (unsigned i = 0, p = 0; i < n; i++) {
      const float a = frand(), b = frand(), c = frand();
      m_hPos[p++] = 2 * (a - 0.5f);
      m_hPos[p++] = 2 * (b - 0.5f);
      m_hPos[p++] = 2 * (c - 0.5f);
      m_hPos[p++] = 1.0f;
    }
  }
  const int* t = getTypes();
  std::vector<int> types(t, t + n);
  m_status = m_engine->impl->set_particles(m_hPos.data(), types.data(), n);
}

void* ParticleSystem::getCudaPosVBO() const {
  float4* p = nullptr;
  int* t = nullptr;
  uint32_t n = 0;
  return (m_engine && m_engine->impl->view(&p, &t, &n) == SMD_OK) ? (void*)p : nullptr;
}

void* ParticleSystem::getCudaTypesVBO() const {
  float4* p = nullptr;
  int* t = nullptr;
  uint32_t n = 0;
  return (m_engine && m_engine->impl->view(&p, &t, &n) == SMD_OK) ? (void*)t : nullptr;
}

void ParticleSystem::dumpGrid() {
  // particleSystem.cpp:1003-1017 (largest cell occupancy), on the host
  const float* p = getArray(POSITION);
  std::vector<unsigned> cnt(m_params.numCells, 0u);
  unsigned mx = 0;
  for (unsigned i = 0; i < m_numParticles; ++i) {
    const int gx = (int)std::floor((p[4 * i] - m_params.worldOrigin.x) / m_params.cellSize.x);
    const int gy = (int)std::floor((p[4 * i + 1] - m_params.worldOrigin.y) / m_params.cellSize.y);
    const int gz = (int)std::floor((p[4 * i + 2] - m_params.worldOrigin.z) / m_params.cellSize.z);
    const unsigned h = (((unsigned)gz & (m_params.gridSize.z - 1)) * m_params.gridSize.y +
                        ((unsigned)gy & (m_params.gridSize.y - 1))) * m_params.gridSize.x +
                       ((unsigned)gx & (m_params.gridSize.x - 1));
    mx = std::max(mx, ++cnt[h]);
  }
  printf("maximum particles per cell = %u\n", mx);
}

void ParticleSystem::dumpParticles(unsigned int start, unsigned int count) {
  const float* p = getArray(POSITION);
  for (unsigned i = start; i < start + count && i < m_numParticles; i++)
    printf("pos: (%.4f, %.4f, %.4f, %.4f)\n", p[i * 4 + 0], p[i * 4 + 1], p[i * 4 + 2], p[i * 4 + 3]);
}

int ParticleSystem::setParams(const SimParams& p) {
  smd_sim_spec s = m_spec;
  for (unsigned g : {p.gridSize.x, p.gridSize.y, p.gridSize.z})
    if (g < 4 || (g & (g - 1))) {
      smd::set_error("gridSize must be a power of two >= 4 per axis");
      return m_status = SMD_ERR_INVALID;
    }
  s.grid_bits[0] = ilog2(p.gridSize.x);
  s.grid_bits[1] = ilog2(p.gridSize.y);
  s.grid_bits[2] = ilog2(p.gridSize.z);
  s.cell_size[0] = p.cellSize.x;
  s.cell_size[1] = p.cellSize.y;
  s.cell_size[2] = p.cellSize.z;
  s.origin[0] = p.worldOrigin.x;
  s.origin[1] = p.worldOrigin.y;
  s.origin[2] = p.worldOrigin.z;
  s.particle_radius = p.particleRadius;
  unsigned n = 0;
  smd_engine_get(m_engine, 0, nullptr, nullptr, 0, &n);
  std::vector<float> pos(4 * (size_t)std::max(n, 1u));
  std::vector<int> types(std::max(n, 1u));
  smd_engine_get(m_engine, 0, pos.data(), types.data(), n, &n);
  const unsigned S = m_engine->impl->tabs[0].numStreams;
  std::vector<uint32_t> st((size_t)S * 20);
  smd_engine_get_streams(m_engine, 0, st.data(), S);
  const smd_sim_spec old = m_spec;
  m_spec = s;
  const int rc = rebuild(pos.data(), types.data(), n, st.data());
  if (rc != SMD_OK) {
    m_spec = old;
    return m_status = rc;
  }
  m_params = p;
  fillReactions();
  return m_status = SMD_OK;
}

std::vector<unsigned int> ParticleSystem::getSpeciesCounts() {
  std::vector<unsigned> c(SMD_MAX_TYPES, 0u);
  if (m_engine) m_status = smd_engine_species_counts(m_engine, 0, c.data());
  c.resize(m_spec.num_types);
  return c;
}

// ---------------------------------------------------------------- C wrappers
struct smd_ps {
  ParticleSystem* impl;
};

extern "C" {

smd_ps* smd_ps_create_reference(unsigned num_particles, unsigned gx, unsigned gy, unsigned gz) {
  ParticleSystem* p = new ParticleSystem(num_particles, smd_u3{gx, gy, gz}, false);
  if (p->lastStatus() != 0) {
    delete p;
    return nullptr;
  }
  return new smd_ps{p};
}

smd_ps* smd_ps_create(const smd_sim_spec* spec, const float* pos4, const int* types, unsigned n, int device) {
  if (!spec) return nullptr;
  ParticleSystem* p = new ParticleSystem(*spec, pos4, types, n, device);
  if (p->lastStatus() != 0) {
    delete p;
    return nullptr;
  }
  return new smd_ps{p};
}

void smd_ps_destroy(smd_ps* ps) {
  if (!ps) return;
  delete ps->impl;
  delete ps;
}

int smd_ps_update(smd_ps* ps) {
  if (!ps) return -SMD_ERR_INVALID;
  const bool ok = ps->impl->update();
  if (ps->impl->lastStatus() != 0) return -ps->impl->lastStatus();
  return ok ? 1 : 0;
}

smd_status smd_ps_reset(smd_ps* ps, int config) {
  if (!ps || config < 0 || config > 1) return SMD_ERR_INVALID;
  ps->impl->reset(config == 0 ? ParticleSystem::CONFIG_RANDOM : ParticleSystem::CONFIG_GRID);
  return (smd_status)ps->impl->lastStatus();
}

int smd_ps_num_particles(smd_ps* ps) { return ps ? ps->impl->getNumParticles() : -SMD_ERR_INVALID; }

smd_status smd_ps_get_positions(smd_ps* ps, float* out4, unsigned cap) {
  if (!ps || !out4) return SMD_ERR_INVALID;
  const float* p = ps->impl->getArray(ParticleSystem::POSITION);
  const unsigned n = std::min<unsigned>(cap, (unsigned)ps->impl->getNumParticles());
  if (p && n) std::memcpy(out4, p, 16 * (size_t)n);
  return (smd_status)ps->impl->lastStatus();
}

smd_status smd_ps_get_types(smd_ps* ps, int* out, unsigned cap) {
  if (!ps || !out) return SMD_ERR_INVALID;
  const int* t = ps->impl->getTypes();
  const unsigned n = std::min<unsigned>(cap, (unsigned)ps->impl->getNumParticles());
  if (t && n) std::memcpy(out, t, 4 * (size_t)n);
  return (smd_status)ps->impl->lastStatus();
}

smd_status smd_ps_set_positions(smd_ps* ps, const float* data4, int start, int count) {
  if (!ps || !data4) return SMD_ERR_INVALID;
  ps->impl->setArray(ParticleSystem::POSITION, data4, start, count);
  return (smd_status)ps->impl->lastStatus();
}

smd_status smd_ps_species_counts(smd_ps* ps, unsigned* counts) {
  if (!ps || !counts) return SMD_ERR_INVALID;
  const std::vector<unsigned> c = ps->impl->getSpeciesCounts();
  std::copy(c.begin(), c.end(), counts);
  return (smd_status)ps->impl->lastStatus();
}

smd_status smd_ps_set_params(smd_ps* ps, const SimParams* params) {
  if (!ps || !params) return SMD_ERR_INVALID;
  return (smd_status)ps->impl->setParams(*params);
}

smd_status smd_ps_get_params(smd_ps* ps, SimParams* params) {
  if (!ps || !params) return SMD_ERR_INVALID;
  *params = ps->impl->getParams();
  return SMD_OK;
}

}  // extern "C"
