/*
 * mercury_hip.h -- C-ABI of the MI355X-native batched mass-spring step.
 *
 * Drop-in boundary for the reference's host simulator API
 * (common/simulator/Simulator.h:19-149 of DawsonCohen/Mercury).  Plain C
 * types only: no torch, no Eigen, no STL.  Every call returns MH_OK (0) or a
 * negative error code; mh_last_error() gives the calling thread's message
 * for its last failing call.  Handles are independent (no process-global
 * simulation state) and each handle is used from one host thread at a time.
 *
 * Reference interface each group replaces (paths in the reference repo):
 *   mh_materials_*     materials::{matLookup, get_composite_id, avg,
 *                      id_lookup, decode, encodedCompositeIdx}
 *                      common/simulator/material.h:109-322
 *   mh_voxel_build     VoxelRobot::Build / Strip / BuildSpringsRecurse /
 *                      ShiftX / ShiftY  common/evolvables/VoxelRobot.cpp:30-331
 *   mh_rest_lengths    SoftBody::Reset spring mean lengths
 *                      common/evolvables/SoftBody.cpp:68-80
 *   mh_baseline        SoftBody::updateBaseline/updateCOM/updateLength
 *                      common/evolvables/SoftBody.h:78-82, SoftBody.cpp:98-122
 *   mh_sim_create      Simulator::Initialize  Simulator.cu:103-198
 *   mh_sim_set_population  Simulator::SetElements / AllocateElement
 *                      Simulator.cu:206-367, 460-529
 *   mh_sim_simulate    Simulator::Simulate  Simulator.cu:369-458 with
 *                      integrateBodies  sim_kernel.cu:271-325
 *   mh_sim_collect     Simulator::Collect  Simulator.cu:537-595
 *   mh_sim_fitness     SoftBody::updateFitness  SoftBody.cpp:124-137
 *   mh_sim_reset / mh_sim_total_time / mh_sim_set_time_step
 *                      Simulator::{Reset, getTotalTime, setTimeStep,
 *                      getDeltaT}  Simulator.h:42-48
 */
#ifndef MERCURY_HIP_H
#define MERCURY_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MH_OK 0
#define MH_ERR_ARG (-1)      /* invalid argument */
#define MH_ERR_HIP (-2)      /* HIP runtime failure (no device, launch) */
#define MH_ERR_RANGE (-3)    /* size beyond what a kernel supports */
#define MH_ERR_STATE (-4)    /* call out of order (e.g. no population) */
#define MH_ERR_NOMEM (-5)    /* host or device allocation failed */
#define MH_ERR_INTERNAL (-6)

/* Message of the calling thread's last failing call ("" if none). */
const char* mh_last_error(void);
int mh_version(void);

/* ------------------------------------------------------------ materials */
typedef struct mh_material_set mh_material_set;
typedef struct {
  float k, dL0, omega, phi;
} mh_base_material;

/* The reference's 22 base materials (+ air), material.h:111-133. */
int mh_materials_reference(mh_material_set** out);
/* n_base non-air base materials; base i (0-based) gets id i+1 and encoding
 * bit i+1; bit 0 is air.  Composites follow the reference rules. */
int mh_materials_create(int n_base, const mh_base_material* base,
                        mh_material_set** out);
int mh_materials_destroy(mh_material_set* set);
/* Composite count = n_base*(n_base+1)/2 + 1 (254 for the reference set). */
int mh_materials_count(const mh_material_set* set);
/* id_lookup(id): (k, dL0, omega, phi) and encoding. */
int mh_materials_by_id(const mh_material_set* set, int id, float kdwp[4],
                       uint32_t* encoding);
/* decode(encoding): composite id, and encodedCompositeIdx(encoding). */
int mh_materials_decode(const mh_material_set* set, uint32_t encoding,
                        uint32_t* id, int32_t* encoded_index);

/* ---------------------------------------------------------- voxel build */
/* Upper bound on the spring count of an nx*ny*nz lattice. */
int mh_voxel_max_springs(int nx, int ny, int nz);
/* voxel_encoding[nx*ny*nz]: material encoding of each voxel (decoded with
 * the set's decode table).  Writes nx*ny*nz masses and the springs in the
 * reference's DFS emission order; *n_springs receives the count.
 * stripped_encoding (may be NULL) receives the voxel encodings after Strip. */
int mh_voxel_build(const mh_material_set* set, int nx, int ny, int nz,
                   float resolution, const uint32_t* voxel_encoding,
                   float* mass_pos, uint8_t* mass_material_id,
                   uint16_t* spring_pairs, float* spring_length,
                   uint8_t* spring_material_id, uint32_t* spring_encoding,
                   int max_springs, int* n_springs,
                   uint32_t* stripped_encoding);
/* VoxelRobot::BatchBuild (VoxelRobot.cpp:291-308): mh_voxel_build for
 * n_robots lattices (voxel_encoding holds n_robots * nx*ny*nz encodings) on
 * `threads` host threads (0 = all), written as one concatenated population:
 * masses n_robots * nx*ny*nz (positions, material ids), springs per robot
 * at [spring_offset[r], spring_offset[r+1]) in each robot's DFS order.
 * spring_capacity bounds the total; mh_voxel_max_springs gives the per-robot
 * bound. */
int mh_voxel_build_batch(const mh_material_set* set, int n_robots, int nx, int ny, int nz,
                         float resolution, const uint32_t* voxel_encoding, int threads,
                         float* mass_pos, uint8_t* mass_material_id, uint32_t* spring_offset,
                         uint16_t* spring_pairs, float* spring_length,
                         uint8_t* spring_material_id, int64_t spring_capacity);
/* No reference counterpart: what mh_voxel_build gives every robot of an
 * nx*ny*nz lattice whatever its materials (VoxelRobot::Build's DFS visits
 * every lattice point, VoxelRobot.cpp:310-331): mass positions (3 per
 * point, shifted), the springs in DFS order and their lengths, and per
 * spring the up to 4 lattice points (0xffff = none) whose voxels decide
 * its material: decode(OR of their non-air encoding bits).  spring_voxels
 * holds 4 entries per spring. */
int mh_voxel_template(int nx, int ny, int nz, float resolution, float* mass_pos,
                      uint16_t* spring_pairs, float* spring_length,
                      uint16_t* spring_voxels, int max_springs, int* n_springs);
/* Mean lengths |p[m0] - p[m1]| as SoftBody::Reset computes them. */
int mh_rest_lengths(const float* pos, const uint16_t* pairs, int n_springs,
                    float* out);
/* Baseline COM (3 floats) and length of one robot (non-air masses). */
int mh_baseline(const float* pos, const uint8_t* mass_material_id,
                int n_masses, float com[3], float* length);

/* ------------------------------------------------------------ simulator */
enum { MH_SOLVER_XPBD_REF = 0, MH_SOLVER_FORCE = 1 };
/* Step kernel families, as a bit set (MH_KERNEL_AUTO = all):
 *   LDS     robots resident in LDS, spring terms gathered per mass;
 *   STREAM  robots over the LDS budget, spring records streamed per substep;
 *   MASS    robots resident in LDS, one lane per mass evaluating each of
 *           its springs (both endpoints evaluate a spring);
 *   MASS_GATHER  robots resident in LDS, one lane per mass; a spring is
 *           evaluated by its first endpoint and its second endpoint reads
 *           the stored term (XPBD solver);
 *   MASS_STREAM  robots over the LDS budget (force solver, up to 2048
 *           masses of at most 255 springs each), one lane per mass
 *           evaluating each of its springs, spring records streamed per
 *           substep.
 * A robot can use the MASS kinds when its masses fit one workgroup (up to
 * 1024), no mass is the first (or the second) endpoint of more than 13 to
 * 16 springs (by workgroup size), and its springs have at most 511
 * distinct (rest length, material) pairs; other robots go to LDS or
 * STREAM.  The kinds allowed never change results, only speed. */
enum { MH_KERNEL_AUTO = 0, MH_KERNEL_LDS = 1, MH_KERNEL_STREAM = 2, MH_KERNEL_MASS = 4,
       MH_KERNEL_MASS_GATHER = 8, MH_KERNEL_MASS_STREAM = 16 };
enum { MH_FITNESS_REFERENCE = 0, MH_FITNESS_DISPLACEMENT = 1 };

typedef struct {
  int solver;          /* MH_SOLVER_* */
  float dt;            /* time step, s */
  float gravity;       /* force solver: m/s^2 along -y */
  float ground_k;      /* force solver: ground stiffness N/m, 0 = no ground */
  float mu_s, mu_k;    /* force solver: static / kinetic friction */
  float drag;          /* force solver: linear drag coefficient N s/m */
  float damping;       /* force solver: velocity factor per step */
  int device;          /* HIP device ordinal */
  int kernel;          /* bit set of MH_KERNEL_* families; 0 = all */
  int block_size;      /* 0 = any; else 256, 512, 768 or 1024 threads */
  int robots_per_block;/* 0 = as many as fit; else an upper bound */
  int substeps_per_launch; /* 0 = tuned (one launch per call until
                          mh_sim_autotune has chosen) */
  int use_graph;       /* capture each simulate's launches in a hipGraph */
} mh_config;

/* Reference defaults: XPBD, dt = 1e-4 (Simulator.h:55), the Land
 * environment constants of environment.h:17-23 for the force solver
 * except ground_k (1e5, stable at dt = 1e-4 with 1 g masses). */
void mh_config_default(mh_config* cfg);

typedef struct {
  int32_t n_robots;
  const uint32_t* mass_offset;        /* n_robots+1 prefix offsets */
  const uint32_t* spring_offset;      /* n_robots+1 prefix offsets */
  const float* pos;                   /* 3*M, xyz per mass */
  const float* vel;                   /* 3*M or NULL (zero) */
  const float* mass;                  /* M or NULL (1.0) */
  const uint8_t* mass_material_id;    /* M composite id; 0 = air (fitness) */
  const uint16_t* spring_pairs;       /* 2*S, robot-local mass indices */
  const float* spring_mean_length;    /* S */
  const uint8_t* spring_material_id;  /* S composite id; 0 = air (skipped) */
} mh_population;

typedef struct mh_sim mh_sim;

int mh_sim_create(const mh_config* cfg, const mh_material_set* set,
                  mh_sim** out);
int mh_sim_destroy(mh_sim* sim);
/* Copies the population to the device (uneven robots allowed), builds the
 * per-block packing, CSR gather lists and material tables, zeroes stresses
 * and the simulated time. */
int mh_sim_set_population(mh_sim* sim, const mh_population* pop);
/* No reference counterpart: Simulator::SetElements for voxel robots of one
 * nx*ny*nz lattice (at most 1024 points), built on the device from their
 * voxel encodings (VoxelRobot::Build per robot, VoxelRobot.cpp:30-331, with
 * the rules mh_voxel_build restates; BatchBuild, VoxelRobot.cpp:291-308).
 * voxel_encoding holds n_robots * nx*ny*nz encodings, host or device
 * memory; nothing else crosses to the device.  Every robot gets the masses
 * and springs mh_voxel_build gives it (positions after ShiftX/ShiftY,
 * `mass` per mass, zero velocity; springs in DFS order, the builder's
 * lengths as mean lengths): simulate, collect (stress per spring in each
 * robot's DFS order, air springs included), fitness and com give the
 * results of mh_voxel_build_batch followed by mh_sim_set_population,
 * bit for bit.  Runs on the mass-parallel kernel, one robot per workgroup
 * (MH_ERR_ARG when the config excludes it); MH_ERR_RANGE when a robot has
 * more than 511 distinct (rest length, material) pairs. */
int mh_sim_set_voxel_population(mh_sim* sim, int n_robots, int nx, int ny, int nz,
                                float resolution, float mass,
                                const uint32_t* voxel_encoding);
/* No reference counterpart: the materials of a device-built voxel
 * population (each may be NULL): mass ids (n_robots * points), spring ids
 * (n_robots * springs, DFS order), voxel encodings after Strip. */
int mh_sim_voxel_materials(mh_sim* sim, uint8_t* mass_material_id,
                           uint8_t* spring_material_id, uint32_t* stripped_encoding);
/* Runs the reference's float time loop: while (remaining > 0) { step(t);
 * t += dt; remaining -= dt; }.  Asynchronous: returns once launched. */
int mh_sim_simulate(mh_sim* sim, float duration, int track_stresses);
/* Runs exactly n_steps steps (same time recurrence). */
int mh_sim_step(mh_sim* sim, int n_steps, int track_stresses);
int mh_sim_synchronize(mh_sim* sim);
/* Caller-allocated outputs, each may be NULL: pos/vel 3*M, stress S
 * (original spring order; air springs report what mh_sim_set_stress gave
 * them, else 0). */
int mh_sim_collect(mh_sim* sim, float* pos, float* vel, float* stress);
/* No reference counterpart: sets the accumulated stress of every spring
 * (S floats, original spring order), which mh_sim_set_population zeroes.
 * Air springs are not stepped, so they keep the value given here.  Devo
 * cycles use it to keep accumulating across a population re-upload, as the
 * reference's stress buffer does across Simulator::Devo calls (cleared only
 * by SetElements, Simulator.cu:355). */
int mh_sim_set_stress(mh_sim* sim, const float* stress);
/* Per-robot fitness on the device: MH_FITNESS_REFERENCE gives
 * max((COMx - COMx0)/L, 0) with the baseline taken at set_population;
 * MH_FITNESS_DISPLACEMENT gives COMx - COMx0. */
int mh_sim_fitness(mh_sim* sim, int mode, float* out);
/* Per-robot COM of non-air masses (3 floats each), device-computed. */
int mh_sim_com(mh_sim* sim, float* out);
int mh_sim_reset(mh_sim* sim);               /* simulated time := 0 */
float mh_sim_total_time(const mh_sim* sim);
/* No reference counterpart: sets the simulated time (the actuation clock),
 * which mh_sim_set_population resets to 0.  Devo cycles use it to continue
 * the clock across a population re-upload, as the reference's in-place
 * spring replacement (Simulator::Devo, Simulator.cu:640-668) does. */
int mh_sim_set_total_time(mh_sim* sim, float t);
float mh_sim_get_time_step(const mh_sim* sim);
int mh_sim_set_time_step(mh_sim* sim, float dt);
/* Population counts and the launch plan chosen for it.  kernel is the bit
 * set of MH_KERNEL_* families in use (robots over the LDS budget run in the
 * streaming kernel, concurrently with the LDS-resident part). */
typedef struct {
  int n_robots, n_masses, n_springs, n_active_springs;
  int kernel, block_size, n_blocks, substeps_per_launch, lds_bytes;
  int max_masses_per_block, max_springs_per_block;
  int n_lds_robots, n_stream_robots, stream_block_size, robots_per_block;
  int n_stream_springs;  /* active springs handled by the streaming kernel */
  int stream_chunk;      /* springs per LDS chunk of the streaming kernel;
                            0 = terms kept in global memory */
  int graph_builds;      /* hipGraphs captured and instantiated so far (a
                            simulate call replays the last one when its
                            step count, stress flag, substeps per launch and
                            plan are unchanged, whatever the clock) */
  int stream_record_bytes; /* bytes per spring record of the mass-parallel
                              streaming kernel (8, or 6 for populations of
                              at most 15 compact materials); 0 when that
                              kernel is not in the plan */
} mh_sim_info;
int mh_sim_get_info(const mh_sim* sim, mh_sim_info* info);
/* Times the candidate launch plans (LDS variant x robots per block, then
 * streaming variant, then substeps per launch: all of a call in one launch,
 * or 2500, 1000 or 250) on the current population for n_steps each and
 * keeps the fastest; results do not depend on the choice.  The choice is
 * cached per (solver, device, the config's kernel, block, robots-per-block,
 * substeps and graph settings, histogram of robot sizes) for the process,
 * and mh_sim_set_population applies a cached choice to a population of the
 * same shape.  Waits for earlier asynchronous work; the simulated state and
 * time are restored afterwards, also when tuning fails. */
int mh_sim_autotune(mh_sim* sim, int n_steps);
/* Device pointers of the state (float4 per mass: xyz + mass / xyz + 0). */
int mh_sim_device_state(mh_sim* sim, void** pos4, void** vel4,
                        void** stream);
/* Average duration (ms) of the step kernel launches of the last simulate,
 * measured with HIP events on the simulator's stream; 0 if not timed. */
int mh_sim_set_timing(mh_sim* sim, int enable);
int mh_sim_last_kernel_ms(const mh_sim* sim, float* avg_ms, int* launches);

/* ------------------------------------------------------------ NN robots */
/* K nearest neighbours of every mass of every robot, on the device
 * (Triangulation::KNN / KNN_CPU, knn.cu:62-190): nbr/dist are
 * [total masses x K], ranked by (distance, index); a zero distance (the mass
 * itself or a coincident mass) counts as +inf.  1 <= K <= 32, every
 * non-empty robot needs more than K masses. */
int mh_knn_batch(int device, int n_robots, const uint32_t* mass_offset,
                 const float* pos, int K, uint16_t* nbr, float* dist);
/* Springs of one robot from its KNN rows (NNRobot::Build,
 * NNRobot.cpp:282-301): edge (i, nbr[i][r]) for every rank r, skipped when
 * shorter than min_length or non-finite; material decode(enc_i | enc_j). */
int mh_knn_springs(const mh_material_set* set, int n, int K,
                   const uint16_t* nbr, const float* dist,
                   const uint32_t* mass_encoding, float min_length,
                   uint16_t* pairs, float* length, uint8_t* material_id,
                   int max_springs, int* n_springs);
/* Material index per mass from a bias-free MLP on its position
 * (NNRobot::forward, NNRobot.h:41-92): ReLU hidden layers, index of the
 * largest output (first maximum).  weights: row-major [out x in] per layer,
 * concatenated; sizes[0] must be 3. */
int mh_nn_materials(int n_layers, const int* sizes, const float* weights,
                    const float* pos, int n, uint8_t* out_index);

/* ------------------------------------------------------- surrogate MLP */
/* No reference counterpart: the network that pre-screens offspring before
 * the optimizer's evaluate call (Evaluator_impl.h:39-107) simulates them.
 * Fully connected sizes[0] -> ... -> sizes[n_layers]; GELU (tanh form) on
 * hidden layers, linear output.  Weights are bf16 on the device with an
 * fp32 master copy; products accumulate in fp32; He-uniform init from seed. */
typedef struct mh_mlp mh_mlp;
int mh_mlp_create(int device, int n_layers, const int* sizes, uint64_t seed,
                  mh_mlp** out);
int mh_mlp_destroy(mh_mlp* net);
int mh_mlp_param_count(const mh_mlp* net);
/* fp32 master parameters, layer by layer: W_l (row-major [out x in]) then
 * b_l.  set_params also resets the Adam moments. */
int mh_mlp_get_params(const mh_mlp* net, float* out);
int mh_mlp_set_params(mh_mlp* net, const float* in);
/* y[batch x out] = net(x[batch x in]); host buffers. */
int mh_mlp_forward(mh_mlp* net, const float* x, int batch, float* y);
/* No reference counterpart; for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops that score offspring every
 * generation): mh_mlp_forward on device buffers x [batch x in] and
 * y [batch x out], enqueued on `stream` (a hipStream_t; NULL = the
 * network's own stream) without waiting, with the same results.  The hidden
 * activations live in a fixed device array of MH_MLP_SCRATCH_FLOATS floats
 * (16 MiB on each device the library runs on, held until the process
 * ends), used in two halves: batch x (widest hidden layer) over
 * MH_MLP_SCRATCH_FLOATS / 2 gives MH_ERR_RANGE and launches nothing.  All
 * calls on one device share the array, so calls that overlap in time there
 * (other streams, other threads) race; training the same network
 * concurrently races too.  Calls no hipMalloc, hipFree or allocating
 * routine: the array's address is resolved by mh_mlp_create. */
#define MH_MLP_SCRATCH_FLOATS (4u << 20)
int mh_mlp_forward_noalloc(mh_mlp* net, const float* x, int batch, float* y,
                           void* stream);
/* Loss = mean 0.5 (y - t)^2 and its gradient w.r.t. the parameters (same
 * layout as get_params), evaluated with the bf16 weights. */
int mh_mlp_loss_grad(mh_mlp* net, const float* x, const float* target,
                     int batch, float* loss, float* grad);
/* Adam (Kingma & Ba 2015; beta 0.9/0.999, eps 1e-8) over n_epochs of
 * shuffled minibatches of the host data set; returns the last epoch's
 * mean loss. */
int mh_mlp_train(mh_mlp* net, const float* x, const float* target, int n,
                 int batch, int n_epochs, float lr, uint64_t seed,
                 float* final_loss);

/* No reference counterpart: a convolutional feature extractor in front of
 * the surrogate MLP, for genomes of 8^3 voxel materials (0 = empty, 1..4;
 * voxel (x, y, z) at x + 8 (y + 8 z)).  conv 3x3x3 5->32 GELU, conv 32->32
 * GELU, 2x2x2 average pool, conv 32->64 GELU, mean over positions (64
 * features), dense 64->256 GELU ->256 GELU ->1.  bf16 weights with an fp32
 * master copy, fp32 accumulation, deterministic reductions; parameters are
 * laid out layer by layer as W [out x in] then b, with conv weights
 * indexed [out][k * C_in + c], k = (dz+1)*9 + (dy+1)*3 + (dx+1). */
typedef struct mh_cnn mh_cnn;
int mh_cnn_create(int device, uint64_t seed, mh_cnn** out);
int mh_cnn_destroy(mh_cnn* net);
int mh_cnn_param_count(const mh_cnn* net);
int mh_cnn_get_params(const mh_cnn* net, float* out);
int mh_cnn_set_params(mh_cnn* net, const float* in);
int mh_cnn_forward(mh_cnn* net, const uint8_t* genomes, int n, float* y);
int mh_cnn_loss_grad(mh_cnn* net, const uint8_t* genomes, const float* target,
                     int n, float* loss, float* grad);
int mh_cnn_train(mh_cnn* net, const uint8_t* genomes, const float* target,
                 int n, int batch, int n_epochs, float lr, uint64_t seed,
                 float* final_loss);

/* ------------------------------------------------ measurement utilities */
/* No reference counterpart: figures for the roofline of the step kernels.
 * LDS read bandwidth of the device (GB/s), from a ds_read_b128 kernel. */
int mh_calibrate_lds(int device, float* gbytes_per_s);
/* Average time (ms) of a plain float4 copy of `bytes` bytes, the known-size
 * reference for scaling hardware byte counters. */
int mh_calibrate_copy(int device, size_t bytes, int reps, float* avg_ms);
/* FP32 VALU issue rate of the device in 10^9 wave64 instructions per second
 * (independent chains, 8 waves per SIMD; kind 0 v_fma_f32, kind 1
 * v_mul_f32, kind 2 v_fma_f32 with twice the chains, kind 3 v_rcp_f32,
 * kind 4 v_rcp_f32 and v_fma_f32 alternating): the peak for kernels bound
 * by vector-ALU issue, and the cost of a transcendental against it. */
int mh_calibrate_valu(int device, int kind, float* giga_instr_per_s);

/* No reference counterpart: checks, bit for bit, the fast arithmetic the
 * step kernels rely on against the compiler's correctly rounded operations,
 * over the float bit patterns [begin, begin + count).  which: 0 hardware
 * sqrt, 1 the kernels' in-range correctly rounded sqrt, 2 hardware
 * reciprocal, 3 x against the next float (every pattern differs; a check
 * of the checker), 4 the kernels' in-range correctly rounded reciprocal;
 * 5 / 6 count inputs whose hardware sqrt lies above / below sqrtf, 7 / 8
 * the same for the hardware reciprocal (positive inputs); 9 the kernels'
 * fast division against a / b on pseudo-random pairs (pattern i of
 * [begin, begin + count) gives one pair; |a| in [2^-79, 2^63) or 0,
 * b in [2^-48, 2^40)); 10 the hardware sqrt with Tuckerman's rounding
 * test (an alternative to the kernels' sqrt, 1) against sqrtf.
 * *mismatches receives the number of differing results. */
int mh_check_arith(int device, int which, uint32_t begin, uint32_t count,
                   uint64_t* mismatches);

#ifdef __cplusplus
}
#endif
#endif /* MERCURY_HIP_H */
