/*
 * gpusmoldyn_amd: C-ABI of the MI355X (gfx950) Brownian-dynamics
 * reaction-diffusion timestep.
 *
 * Two groups of entry points live here:
 *
 *  1. The simulation API (smd_*): a ParticleSystem-style object over one or
 *     more independent replicas whose particle state stays resident in HBM.
 *     It replaces ParticleSystem (particleSystem.h:50-218): construction
 *     (_initialize, particleSystem.cpp:573-751), update() (:829-1001),
 *     getArray()/setArray() (:1031-1080) and the reaction-table setup of
 *     smolreact.cpp:173-483 and rxnparam.cpp:42-308.
 *
 *  2. The reference's own extern "C" layer (particleSystem.cuh:29-110,
 *     MersenneTwister.h:49, radixsort.h:45-95), with HIP types, declared in
 *     gpusmoldyn_legacy.h.
 *
 * Every function returns an smd_status (0 = success) unless stated
 * otherwise; smd_last_error() gives a message for the last failure on the
 * calling thread.  No function aborts the process.
 */
#ifndef GPUSMOLDYN_H_INCLUDED
#define GPUSMOLDYN_H_INCLUDED

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SMD_MAX_TYPES 16
#define SMD_MAX_RXN 32
#define SMD_MAX_SPHERES 4
#define SMD_MT_PARAM_SETS 32768
#define SMD_GAUSS_DIM 4096

typedef enum smd_status {
  SMD_OK = 0,
  SMD_ERR_INVALID = 1,      /* bad argument or inconsistent specification */
  SMD_ERR_HIP = 2,          /* HIP runtime error */
  SMD_ERR_NO_DEVICE = 3,    /* no GPU visible */
  SMD_ERR_IO = 4,           /* MersenneTwister.dat not found or short */
  SMD_ERR_CAPACITY = 5,     /* particle count exceeds the allowed maximum */
  SMD_ERR_EXTINCT = 6,      /* no particles left (update() returns false) */
  SMD_ERR_UNSUPPORTED = 7,  /* OpenGL interop entry points */
  SMD_ERR_RXN_SETUP = 8     /* rxnsetrate/rxnsetproduct error code, see smd_last_error */
} smd_status;

/* Boundary behaviour of a box face. */
typedef enum smd_boundary {
  SMD_BC_CLAMP = 0,    /* adjustPositions (particles_kernel.cu:60-95): clamp to [lo + r, hi - r] */
  SMD_BC_REFLECT = 1,
  SMD_BC_ABSORB = 2,
  SMD_BC_PERIODIC = 3
} smd_boundary;

/* One reaction (the fields of struct Reaction, reaction.h:36-48, that are
 * inputs; prob/bindRadius/unbindRadius/productPositions are derived). */
typedef struct smd_reaction {
  int order;          /* 0, 1 or 2 */
  int reactants[2];   /* species ids >= 1; -1 when unused */
  int products[2];    /* species ids >= 1; -1 when unused */
  int nprod;          /* 0, 1 or 2 */
  float rate;         /* order 0: molecules/volume/time, 1: 1/time, 2: volume/time */
  float pgem;         /* geminate probability for the unbinding radius of a
                         reversible split; <= 0 selects 0.2 (smolreact.cpp:375) */
  float prob;         /* order 2: reaction probability at contact; <= 0 selects 1
                         (smolreact.cpp:382-383) */
} smd_reaction;

/* Everything that defines one simulation (one replica). */
typedef struct smd_sim_spec {
  int num_types;                 /* species + the dummy type 0 */
  float D[SMD_MAX_TYPES];        /* diffusion coefficients, D[0] = 0 */
  float dt;
  float volume;                  /* system volume (zeroth-order reactions) */
  int num_reactions;
  smd_reaction rxn[SMD_MAX_RXN];
  float box_lo[3], box_hi[3];
  int face[6];                   /* smd_boundary per face: -x,+x,-y,+y,-z,+z */
  float particle_radius;         /* clamp margin (particleSystem.cpp:69) */
  int num_spheres;
  float sphere[SMD_MAX_SPHERES][4];  /* centre x,y,z and radius; reflective from both sides */
  int unbind_offsets;            /* 1: order-1 splits place products at the rotated
                                    unbinding offsets; 0: both at the reaction point
                                    (reaction_kernel.cuh:320-321) */
  unsigned grid_bits[3];         /* log2 of the hash grid dimension per axis (>= 2) */
  float cell_size[3];            /* <= 0: (box_hi - box_lo) / 2^grid_bits */
  float origin[3];               /* hash grid origin (usually box_lo) */
  unsigned num_streams;          /* S: particle i draws from stream i mod S (0 -> 32768) */
  unsigned seed;                 /* master seed of the stream seeds (4357) */
} smd_sim_spec;

/* Derived per-reaction parameters (what rxnsetrate/rxnsetproduct compute). */
typedef struct smd_rxn_derived {
  float prob;
  float bind_radius2;      /* bindRadiusSquared (reaction.h:39) */
  float bind_radius;       /* sqrt of bindRadiusSquared */
  float unbind_radius;
  float product_pos[6];
  int setup_code;          /* rxnsetrate/rxnsetproduct return code (0 ok) */
} smd_rxn_derived;

typedef struct smd_engine smd_engine;

const char* smd_last_error(void);
int smd_version(void);

/* Path of MersenneTwister.dat (32768 16-byte records).  Default: the
 * package's data directory, or $SMD_MT_DAT. */
smd_status smd_set_mt_dat_path(const char* path);

/* ---- host-side setup (no GPU needed) ---- */

/* createGaussTable (stats.cpp:277-282): 4096 standard-normal quantiles,
 * mirrored, then passed through the reference's shuffle. */
smd_status smd_gauss_table(float* out, int n);
/* rxnparam.cpp:42-308 (FP64). */
double smd_numrxnrate(double step, double a, double b);
double smd_actrxnrate(double step, double a);
double smd_bindingradius(double rate, double dt, double difc, double b, int rel);
double smd_unbindingradius(double pgem, double dt, double difc, double a);
/* Radial distribution function of a reactive pair and the rate tables of
 * rxnparam (rxnparam.cpp:325-564), FP64, lengths reduced by the binding
 * radius.  smd_rdf_steady_state iterates diffusion (rms step `step`), the
 * reverse-reaction source at b (b >= 0; b < 0: irreversible) and absorption
 * inside r = 1 to a steady flux (relative change < eps); rdfa is the trial rdf
 * in and the absorbed rdf out, rdfd the diffused rdf out; returns the flux per
 * step (reduced rate), or -1.  smd_rdf_sweep reproduces one column of the
 * table maker (rdfmaketable, rxnparam.cpp:477-564): mode 'i' (irreversible),
 * 'r' (b >= 1) or 'b' (0 <= b <= 1), n radial points, the 31 reduced steps
 * exp(-3) .. exp(3) swept upwards (down = 0) or downwards, each steady state
 * starting from the previous; flux31[k] belongs to step exp(-3) exp(0.2)^k. */
double smd_rdf_steady_state(const double* r, double* rdfa, double* rdfd, int n, double step, double b, double eps);
smd_status smd_rdf_sweep(char mode, int n, double eps, int down, double b, double* flux31);
/* The radial grid of the table maker: 0, rmax/n, ..., with the points next
 * to 1 moved to 0.9999 and 1.0001 (n points). */
smd_status smd_rdf_grid(double rmax, int n, double* r);
/* Many sweeps at once on GPU `device` (one workgroup per sweep, n <= 1024):
 * sweep q has mode modes[q], unbinding radius bs[q] and direction downs[q];
 * flux[31 q + k] belongs to reduced step exp(-3) exp(0.2)^k. */
smd_status smd_rdf_sweeps_gpu(int device, int n, double eps, int n_sweeps, const char* modes, const double* bs,
                              const int* downs, double* flux);

/* Derived reaction parameters of a spec (smolreact.cpp:173-483). */
smd_status smd_spec_derive(const smd_sim_spec* spec, smd_rxn_derived* out);
/* Hash-grid cell size and origin the engine uses for a spec (periodic axes
 * get a cell size that tiles the box exactly). */
smd_status smd_spec_cells(const smd_sim_spec* spec, float* cell3, float* origin3);
/* Seeds of S streams: output k of the dcmt stream with parameter set 0
 * seeded with `seed` (MersenneTwister.dat record 0). */
smd_status smd_stream_seeds(unsigned seed, unsigned count, uint32_t* out);
/* The dcmt parameters (matrix_a, mask_b, mask_c) of record `index`. */
smd_status smd_mt_params(unsigned index, uint32_t* abc);
/* `count` draws of unirandOOF(lo, hi) (stats.h:35-37) from the SFMT-19937
 * generator after init_gen_rand(seed): the sequence behind
 * ParticleSystem::randPos (particleSystem.h:115-120). */
smd_status smd_sfmt_uniform(unsigned seed, unsigned count, float lo, float hi, float* out);

/* ---- simulation engine ---- */

smd_status smd_device_count(int* n);
/* Create an engine on GPU `device`.  radix_bits: sort digit width 4, 8 or
 * 11; 0 = automatic (8).  The sorted order, hence every result, is the same
 * for any width. */
smd_engine* smd_engine_create(int device, int radix_bits);
void smd_engine_destroy(smd_engine* e);
/* Add a replica with n initial particles (pos: n float4, types: n ints).
 * Returns the replica id (>= 0) or -smd_status. */
int smd_engine_add_replica(smd_engine* e, const smd_sim_spec* spec, const float* pos4, const int* types, unsigned n);
/* Allocate device memory for `capacity` particles in total (0: automatic),
 * upload, initialise the random streams. */
smd_status smd_engine_finalize(smd_engine* e, unsigned capacity);
/* Advance every replica by `nsteps` timesteps.  Asynchronous with respect to
 * the host except when capacity must be checked. */
smd_status smd_engine_step(smd_engine* e, int nsteps);
/* Replay pairs of steps as one captured hipGraph (default on); off = one
 * launch sequence per step. */
smd_status smd_engine_set_graphs(smd_engine* e, int on);
/* Launch shapes of the step kernels, for the autotuner: every setting gives
 * the same results bit for bit.  SMD_KNOB_PRODUCER_BLOCK: threads per
 * workgroup of the diffuse/react/hash kernel (128, 256 = default, 512; one
 * thread per random stream); SMD_KNOB_PAIR_BLOCK: of the pair search (64 =
 * default, 128, 256; one thread per sorted particle).  No counterpart in
 * the reference (fixed 256-thread diffuse and 64-thread collide blocks,
 * particleSystemCuda.cu:167, 336). */
typedef enum smd_knob {
  SMD_KNOB_PRODUCER_BLOCK = 0,
  SMD_KNOB_PAIR_BLOCK = 1,
  /* 1 (default): an engine of several replicas sorts each replica's keys on
   * its local cells (two digit passes where the global keys need three);
   * 0: one sort over all keys */
  SMD_KNOB_SEGMENTED_SORT = 2
} smd_knob;
smd_status smd_engine_set_knob(smd_engine* e, int knob, int value);
/* Current value of a knob, or -SMD_ERR_INVALID. */
int smd_engine_get_knob(const smd_engine* e, int knob);
/* Key sort of the step: 1 = count, scan and scatter kernels per digit
 * (default), 0 = single sweep (one histogram kernel, one binning kernel per
 * digit with decoupled look-back).  Both give the same stable order. */
smd_status smd_engine_set_sort(smd_engine* e, int kind);
smd_status smd_engine_sync(smd_engine* e);
/* Current particle count of each replica (synchronises). */
smd_status smd_engine_counts(smd_engine* e, unsigned* counts);
unsigned smd_engine_num_replicas(const smd_engine* e);
/* Sort digit width in use (fixed at finalize when created with 0). */
int smd_engine_radix_bits(const smd_engine* e);
/* Sum over all steps so far of the particles entering each step (all
 * replicas): the numerator of particle-steps/s. */
smd_status smd_engine_particle_steps(smd_engine* e, unsigned long long* out);
/* Copy replica r's particles to the host (at most cap). *n_out = count. */
smd_status smd_engine_get(smd_engine* e, int r, float* pos4, int* types, unsigned cap, unsigned* n_out);
/* Species histogram of replica r (counts[t] for t < num_types). */
smd_status smd_engine_species_counts(smd_engine* e, int r, unsigned* counts);
/* RNG stream state of replica r: 20 words per stream, stream-major
 * (19 state words then iState). */
smd_status smd_engine_get_streams(smd_engine* e, int r, uint32_t* words, unsigned cap_streams);
smd_status smd_engine_set_streams(smd_engine* e, int r, const uint32_t* words, unsigned n_streams);
/* Derived reaction parameters of replica r. */
smd_status smd_engine_rxn(smd_engine* e, int r, smd_rxn_derived* out);
/* Stage-level hooks used by the parity tests: run only the producer (or the
 * producer + sort + cell table + candidate search) of one step and copy the
 * intermediate arrays out.  stage: 1 = producer, 2 = + sort, 3 = + candidates,
 * 4 = + resolve.  Arrays are sized for the current total particle count. */
smd_status smd_engine_debug_stage(smd_engine* e, int stage, float* pos4, int* types, uint32_t* keys_sorted,
                                  uint32_t* vals_sorted, int* cand, uint32_t* n_total, uint32_t* n_live);
/* After smd_engine_debug_stage with stage >= 2: the cell table, first[h] =
 * lower bound of key h in the sorted keys for h <= total buckets (key total
 * buckets marks removed particles, so first[total buckets] = live count) and
 * first[total buckets + 1] = particle count: the cellStart/cellEnd of
 * particles_kernel.cu:209-272 for every bucket.  In a slot engine the
 * entries of a slot past its job's grid (offset 2^(grid bits) + 1 onward, or
 * past offset 0 for a slot never used) are not maintained: no key falls there.
 * cap words at most; *n_out = total buckets + 2. */
smd_status smd_engine_debug_first(smd_engine* e, uint32_t* first, uint32_t cap, uint32_t* n_out);
/* Number of buckets (keys deadKey = total buckets) and key bits. */
smd_status smd_engine_layout(smd_engine* e, uint32_t* total_buckets, uint32_t* key_bits, uint32_t* total_streams,
                             uint32_t* capacity);
/* How the step sorts its keys: *passes digit passes, *segmented = 1 when each
 * replica is sorted on its local cells (SMD_KNOB_SEGMENTED_SORT applies),
 * 0 for one sort over the global keys.  No counterpart in the reference. */
smd_status smd_engine_sort_plan(const smd_engine* e, int* passes, int* segmented);
/* Per-kernel timing of the last timed step (HIP events, microseconds):
 * names is a '\n' separated list written into buf. */
smd_status smd_engine_profile_step(smd_engine* e, int nsteps, double* us, int max_stages, char* names, int names_cap,
                                   int* n_stages);

/* Serving (no counterpart in the reference, whose host loop runs one
 * simulation per process): an engine of `n_slots` empty replica slots, each
 * owning 2^slot_bucket_bits hash buckets and `slot_streams` RNG streams.
 * Jobs are admitted into free slots between steps and are stepped together
 * with every other resident job; a job's results do not depend on what else
 * is resident.  inject_cap bounds the particles admitted between two steps.
 * Call on a fresh engine instead of add_replica + finalize. */
smd_status smd_engine_finalize_slots(smd_engine* e, unsigned n_slots, unsigned slot_bucket_bits,
                                     unsigned slot_streams, unsigned capacity, unsigned inject_cap);
/* Load a job into an empty slot: its tables, its streams (seeded from
 * spec->seed) and its n particles, which enter at the next step's
 * compaction (the job's first step is the step after that). */
smd_status smd_engine_admit(smd_engine* e, int slot, const smd_sim_spec* spec, const float* pos4, const int* types,
                            unsigned n);
/* Remove every particle of a slot during the next step; the slot is empty
 * (and may be admitted again) once that step has been enqueued. */
smd_status smd_engine_retire(smd_engine* e, int slot);

/* Draw `count` numbers from each of `n_streams` fresh streams initialised
 * with (params record k % 32768, seeds[k]) on the GPU: out[k * count + j]. */
smd_status smd_mt_generate(int device, const uint32_t* seeds, unsigned n_streams, int count, int batch3,
                           uint32_t* out);

/* Stable sort of (key, value) pairs on the GPU by the low key_bits bits
 * (RadixSort::sort, radixsort.h:168-185).  Host arrays, n elements;
 * radix_bits 4, 8 or 11. */
smd_status smd_sort_pairs(int device, uint32_t* keys, uint32_t* values, unsigned n, unsigned key_bits,
                          int radix_bits);

/* Measurement helper (no counterpart in the reference): times `reps`
 * launches of a plain float4 copy kernel over `bytes` bytes (read once,
 * written once) on `device`; *us = average launch duration.  Used to scale
 * hardware byte counters (FETCH_SIZE, WRITE_SIZE) to bytes moved. */
smd_status smd_copy_probe(int device, unsigned long long bytes, int reps, double* us);

/* ---- ParticleSystem (particleSystem.h:50-218) through C ----
 * The C++ class is declared in particle_system.h; these wrappers expose it to
 * C and foreign-function callers.  smd_ps_update returns 1 (true), 0 (false:
 * extinction or more than MAX_PARTICLES) or a negative status. */
typedef struct smd_ps smd_ps;
struct SimParams;
/* The reference constructor: ParticleSystem(numParticles, gridSize, false). */
smd_ps* smd_ps_create_reference(unsigned num_particles, unsigned gx, unsigned gy, unsigned gz);
smd_ps* smd_ps_create(const smd_sim_spec* spec, const float* pos4, const int* types, unsigned n, int device);
void smd_ps_destroy(smd_ps* ps);
int smd_ps_update(smd_ps* ps);
/* config: 0 = CONFIG_RANDOM, 1 = CONFIG_GRID */
smd_status smd_ps_reset(smd_ps* ps, int config);
int smd_ps_num_particles(smd_ps* ps);
/* getArray(POSITION): copies numParticles float4 (at most cap) */
smd_status smd_ps_get_positions(smd_ps* ps, float* out4, unsigned cap);
smd_status smd_ps_get_types(smd_ps* ps, int* out, unsigned cap);
/* setArray(POSITION, data, start, count) */
smd_status smd_ps_set_positions(smd_ps* ps, const float* data4, int start, int count);
smd_status smd_ps_species_counts(smd_ps* ps, unsigned* counts);
smd_status smd_ps_set_params(smd_ps* ps, const struct SimParams* params);
smd_status smd_ps_get_params(smd_ps* ps, struct SimParams* params);

#ifdef __cplusplus
}
#endif

#endif /* GPUSMOLDYN_H_INCLUDED */
