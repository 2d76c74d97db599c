/*
 * CPU oracle of the gpusmoldyn timestep -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain C++ restatement of the reference's per-step algorithm for one
 * simulation (one replica), used by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg as the checker.  Nothing in the product links,
 * loads or calls it.
 *
 * Followed reference code (files under the reference tree):
 *   MersenneTwister.cuh:40-77      stream init and draw
 *   rand_kernel.cuh:32-51          randReal / thetarandCCF / unirandCOF / coinrandF
 *   particles_kernel.cu:60-95      adjustPositions (clamp)
 *   particles_kernel.cu:125-166    diffuse
 *   particles_kernel.cu:169-207    calcGridPos / calcGridHash / calcHashD
 *   particles_kernel.cu:209-272    reorderDataAndFindCellStartD
 *   particles_kernel.cu:316-497    collideCell / collideD (27-cell scan)
 *   reaction_kernel.cuh:166-258    dotMV / DirCosM / bimReact
 *   reaction_kernel.cuh:263-342    uniReact
 *   particles_kernel.cu:97-123     compactPosAndTypeD (order-preserving intent)
 *   particleSystem.cpp:829-1001    update(): stage order, births appended after survivors
 *   reaction_cpu.cpp:36-63         zeroReact (order-0 creation, appended after the births)
 *   stats.cpp:234-275              poisrandF (multiplication method)
 *   radixsort.cu (semantics only)  stable sort of (hash, index) by hash
 *   surfaces_kernel.cuh:97-248     sphere sidedness / crossing semantics
 *
 * Rules where the reference races or is undefined (the GPU build follows the
 * same rules):
 *   - particle i draws from stream i mod S, streams consumed in increasing i;
 *     per particle: 3 diffusion draws, then the unimolecular coin flips;
 *   - in the pair step, sorted particle j draws from stream j mod S;
 *   - each particle reacts bimolecularly at most once: j and c react when each
 *     is the other's nearest reactive partner (ties: lower sorted index); the
 *     lower sorted index plays mol1 and makes the draws;
 *   - sin/cos/acos of the product rotation use the fixed evaluations below.
 *
 * Parity pins: MT streams against the reference's compiled dcmt generator
 * (rnd/genmtrand.cpp) fixtures; pair candidates against O(N^2) brute force;
 * sort against std::stable_sort semantics; diffusion moments against the
 * Gaussian table's closed form; sin/cos/theta against libm.
 */
#ifndef SMD_ORACLE_H
#define SMD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_TYPES 16
#define ORC_MAX_RXN 32
#define ORC_MAX_SPHERES 4

typedef struct orc_net {
  int num_types, num_rxn;
  float rate[ORC_MAX_TYPES];
  float D[ORC_MAX_TYPES];
  int uni_count[ORC_MAX_TYPES], uni_start[ORC_MAX_TYPES], uni_list[ORC_MAX_RXN];
  float prob[ORC_MAX_RXN], bind_r2[ORC_MAX_RXN];
  int prod_x[ORC_MAX_RXN], prod_y[ORC_MAX_RXN];
  float prod_pos[ORC_MAX_RXN][6];
  int pair_rxn[ORC_MAX_TYPES * ORC_MAX_TYPES];
  int has_uni, has_bim;
  float lo[3], hi[3];
  int face[6];
  float particle_radius;
  int num_spheres;
  float sphere[ORC_MAX_SPHERES][4];
  int unbind_offsets;
  float origin[3], cell[3];
  unsigned gbits[3];
  unsigned num_streams;
  /* order-0 reactions in index order: Poisson(mean) creations per step,
   * mean = prob[zero_list[k]], drawn as zero_pieces[k] Poisson variates by
   * the multiplication method against zero_g[k] = exp(-mean / pieces) */
  int num_zero;
  int zero_list[ORC_MAX_RXN];
  int zero_pieces[ORC_MAX_RXN];
  float zero_g[ORC_MAX_RXN];
} orc_net;

/* One simulation.  pos: cap*4 floats, type: cap ints, mt: num_streams*20
 * words (19 state words + iState per stream), par: num_streams*3 words. */
typedef struct orc_sim {
  orc_net net;
  unsigned n, cap;
  float* pos;
  int* type;
  uint32_t* mt;
  const uint32_t* par;
  uint32_t* zmt;        /* creation stream (20 words), used when num_zero > 0 */
  const uint32_t* zpar; /* its 3 parameter words */
} orc_sim;

/* Stream init (MersenneTwisterInitialise) and draws. */
void orc_mt_init(uint32_t* st20, uint32_t seed);
uint32_t orc_mt_draw(uint32_t* st20, const uint32_t* par3);

void orc_sincos(float x, float* s, float* c);
float orc_theta(float u);

/* One or more timesteps.  Returns 0, or 5 if the capacity would be exceeded. */
int orc_step(orc_sim* s, const float* gauss, int nsteps);

/* Intermediate results of one step (for stage-level parity); the sim is
 * advanced through `stage` (1 producer, 2 sort, 3 candidates, 4 resolve) and
 * left uncompacted.  keys: n (stage 1: unsorted keys, else sorted), vals: n,
 * cand: n_live. */
int orc_step_stage(orc_sim* s, const float* gauss, int stage, uint32_t* keys, uint32_t* vals, int* cand,
                   unsigned* n_live);

/* Pair candidates by O(N^2) brute force over sorted positions/types
 * (minimum image on periodic axes). */
void orc_candidates_bruteforce(const orc_net* net, const float* sorted_pt, unsigned n_live, int* cand);

int orc_num_threads(void);
void orc_set_threads(int n);  // OpenMP threads of the later calls (n > 0)

#ifdef __cplusplus
}
#endif

#endif
