/*
 * oracle.h -- CPU restatement of the Mercury batched mass-spring step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (mercury_amd/, include/)
 * may include, link or call anything declared here.  Only tests/, the
 * smoke() check in __graft_entry__.py and the cpu_baseline leg of bench.py
 * use it, and only as the checker or the timed CPU baseline.
 *
 * Contents and the reference code each part restates (paths relative to the
 * reference repository):
 *   materials  common/simulator/material.h:109-322 (base table, composite
 *              ids, avg, id_lookup, decode, encodedCompositeIdx)
 *   voxel      common/evolvables/VoxelRobot.cpp:30-381 (Strip, DFS spring
 *              build, material rules, ShiftX/ShiftY)
 *   fitness    common/evolvables/SoftBody.cpp:98-137
 *   xpbd       common/simulator/sim_kernel.cu:141-269 (preSolve,
 *              solveDistance, update) as FP32 mirror and FP64 oracle
 *   force      paper Alg. 1 plus the project's land/water spec (no reference
 *              code exists; see the spec comment in step.c)
 *   sin        the project's own float sine (Cody-Waite reduction + Taylor
 *              polynomial), an independent implementation of the same spec
 *              the HIP kernels follow
 *   knn/nn     common/evolvables/knn.cu:147-190 (KNN_CPU), NNRobot.h:41-92
 *   mlp        FP64 forward/backward/Adam of the surrogate MLP
 *
 * Parity pinning of each part is listed in the header of its .c file.
 */
#ifndef MERCURY_ORACLE_H
#define MERCURY_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- materials */
typedef struct {
  uint32_t id;       /* composite id (material.h get_composite_id numbering) */
  float k, dL0, omega, phi;
  uint32_t encoding; /* bit per base material, bit 0 = air */
} orc_material;

/* Reference base set: 23 materials, index 0 = air (material.h:111-133). */
int orc_reference_base(orc_material* out /* [23] */);

/* Composite tables for a base set of n_total materials (index 0 = air).
 * by_id[cc] is id_lookup, by_enc[cc] is the decode table, with
 * cc = (n_total-1)*n_total/2 + 1.  Returns cc. */
int orc_composite_tables(const orc_material* base, int n_total,
                         orc_material* by_id, orc_material* by_enc);
int orc_encoded_idx(uint32_t encoding, int composite_count);
uint32_t orc_composite_id(int n_total, uint32_t id1, uint32_t id2);

/* ------------------------------------------------------------ voxel builder */
/* voxel_enc[nx*ny*nz]: encoding of each voxel's material (decoded with the
 * decode table).  Outputs: n = nx*ny*nz masses (pos xyz after ShiftX/ShiftY,
 * composite id of the mass material), springs in DFS emission order.
 * Returns the spring count, or -1 if max_springs is too small. */
int orc_voxel_build(int nx, int ny, int nz, float resolution,
                    const uint32_t* voxel_enc, const orc_material* by_enc,
                    int composite_count, float* mass_pos, uint32_t* mass_id,
                    uint16_t* pairs, float* spring_len, uint32_t* spring_id,
                    uint32_t* spring_enc, int max_springs,
                    uint32_t* stripped_enc /* may be NULL */);

/* ------------------------------------------------------------------ fitness */
/* pos: 3*n, is_air: n.  Incremental-mean COM over non-air masses
 * (SoftBody.cpp:112-122). */
void orc_com(const float* pos, const uint8_t* is_air, int n, float* com3);
float orc_length(const float* pos, const uint8_t* is_air, int n);
float orc_fitness(float com_x, float base_com_x, float length);

/* ------------------------------------------------------------------ sine */
float orc_sinf(float x);

/* ------------------------------------------------------------- population */
typedef struct {
  int n_robots;
  const uint32_t* mass_off;   /* n_robots+1 */
  const uint32_t* spring_off; /* n_robots+1 */
  const float* mass;          /* M   (force mode) */
  const uint16_t* pairs;      /* 2*S robot-local mass indices */
  const float* lbar;          /* S   mean length */
  const uint8_t* mat;         /* S   composite id into the material table */
} orc_topology;

typedef struct {
  int n;                      /* table length (composite count) */
  const float* k;
  const float* dL0;
  const float* omega;
  const float* phi;
} orc_mattab;

typedef struct {
  float dt;
  float gravity;
  float ground_k;   /* 0 disables the ground plane */
  float mu_s, mu_k;
  float drag;       /* linear drag coefficient c (force = -c v) */
  float damping;    /* velocity factor applied every step */
} orc_env;

/* xpbd_ref: the reference XPBD step.  pos/vel are 3*M, stress S or NULL. */
int orc_xpbd_f32(const orc_topology* topo, const orc_mattab* mt, float dt,
                 float t0, int nsteps, float* pos, float* vel, float* stress,
                 int nthreads);
int orc_xpbd_f64(const orc_topology* topo, const orc_mattab* mt, float dt,
                 float t0, int nsteps, double* pos, double* vel,
                 double* stress, int nthreads);

int orc_force_f32(const orc_topology* topo, const orc_mattab* mt,
                  const orc_env* env, float t0, int nsteps, float* pos,
                  float* vel, int nthreads);
int orc_force_f64(const orc_topology* topo, const orc_mattab* mt,
                  const orc_env* env, float t0, int nsteps, double* pos,
                  double* vel, int nthreads);

/* Step count and start time of every step for Simulate(duration) starting at
 * simulated time t0 (Simulator.cu:428-449).  times may be NULL. */
int orc_step_times(float t0, float dt, float duration, float* times,
                   int max_steps);

/* --------------------------------------------------------------- KNN / NN */
/* KNN_CPU (knn.cu:147-190): for every mass, its K nearest other masses by
 * float distance, ascending, ties by index; zero distance counts as +inf. */
int orc_knn(const float* pos, int n, int K, uint16_t* nbr, float* dist);

/* ------------------------------------------------------------- surrogate */
/* Dense MLP, row-major weights W_l [out x in], biases b_l [out].
 * act: 0 = GELU (tanh form) on hidden layers, linear output.
 * x [batch x in0] -> y [batch x out_last].  Double precision; with
 * bf16_operands the activations and gradients entering each product are
 * rounded to bf16 (via float, nearest even), as the device GEMMs do. */
int orc_mlp_forward(int n_layers, const int* sizes, const double* const* W,
                    const double* const* b, int batch, const double* x,
                    double* y, int bf16_operands);
/* Loss = mean over batch of 0.5*(y - target)^2; gradients written to gW/gb. */
double orc_mlp_loss_grad(int n_layers, const int* sizes,
                         const double* const* W, const double* const* b,
                         int batch, const double* x, const double* target,
                         double* const* gW, double* const* gb, int bf16_operands);
/* One Adam step (Kingma & Ba 2015, Algorithm 1) on n parameters. */
void orc_adam(int n, double* p, const double* g, double* m, double* v,
              double lr, double beta1, double beta2, double eps, int t);

#ifdef __cplusplus
}
#endif
#endif
