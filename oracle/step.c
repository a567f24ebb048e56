/*
 * step.c -- oracle restatement of the batched step.  TEST INFRASTRUCTURE
 * ONLY (see oracle.h).
 *
 * xpbd_ref follows common/simulator/sim_kernel.cu:
 *   E1 preSolve      141-153  x~ = x + v*dt                (float, rn)
 *   E2 solveDistance 166-253  per non-air spring (matId != 0, line 211):
 *        alpha = ((1/k)/dt)/dt                        (220)
 *        r     = dL0 * sin(fma(omega, t, phi))        (222; nvcc contracts)
 *        L     = fma(Lbar, r, Lbar)                   (223)
 *        K     = 2 + alpha                            (225)
 *        d     = |x~a - x~b|,  n = (x~a - x~b)/(d + 1e-12f)   (226-228)
 *        lambda = -(d - L)/K,  dp = lambda*n          (230-232)
 *        stress += lambda/Lbar                        (234)
 *        +dp on a, -dp on b                           (236-242)
 *   E3 x_new = x~ + sum(dp)                           (246-252)
 *   E4 update        255-269  v = (float)(0.99*(double)(x_new - x)/dt)
 *   time: t is the float Simulator::m_total_time before its increment
 *   (common/simulator/Simulator.cu:428-449).
 * The reference sums corrections with shared-memory float atomics (order
 * undefined).  The canonical order here, and in the HIP kernels, is: per
 * mass, one sum from +0.0f over every spring it belongs to, in ascending
 * spring index, adding the spring's term where the mass is the first
 * endpoint and subtracting it where it is the second (first before second
 * for a spring joining a mass to itself).  The term seen from the second
 * endpoint, computed from its own side, is the exact IEEE negation, since
 * a - b == -(b - a) and the length is symmetric.  |.| is
 * sqrtf(fmaf(z, z, fmaf(y, y, x*x))) and sin is orc_sinf (CUDA's
 * norm3df/sinf are outside the reference tree).
 *
 * force (project spec; the reference snapshot has no force solver -- paper
 * p.2 Alg. 1 and the land/water environment of paper p.2 Sec. II.E):
 *   per non-air spring: L and d as above, D = x~a - x~b;
 *        f = (k * (L - d)) * (1 / (d + 1e-12f))
 *        (Hooke's law with one correctly rounded reciprocal per spring);
 *        the term f D enters a's sum as sum = fmaf(f, D, sum) and b's as
 *        sum = fmaf(-f, D, sum), per component, one rounding each
 *        (canonical order above; from b's own side, D is -D exactly)
 *   per mass: F = that sum
 *        F.y = F.y - m*g
 *        F   = F - drag*v                       (if drag != 0)
 *        if ground_k != 0 and y < 0:
 *            Fn = ground_k * (-y);  F.y = F.y + Fn
 *            Fh = sqrtf(F.x*F.x + F.z*F.z)
 *            if Fh <= mu_s*Fn:  F.x = F.z = 0, v.x = v.z = 0     (static)
 *            else: s = 1 - (mu_k*Fn)/Fh;  F.x *= s, F.z *= s     (kinetic)
 *        u = v + (F * (1/m)) * dt;  x = x + u*dt;  v = damping * u
 * The FP64 versions run the same steps in double on the float inputs, with
 * the same float time sequence, using libm sin.
 * Pinned by: closed forms in tests/test_oracle_step.py (single spring
 * lambda, momentum invariant, air no-op, free fall y_n = y0 - g dt^2
 * n(n+1)/2, static hold, kinetic reduction, linear drag decay, two-mass
 * oscillation frequency), FP32 vs FP64 agreement, and brute-force scatter
 * vs gather on small robots.
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define ORC_EPS 1e-12f

int orc_step_times(float t0, float dt, float duration, float* times,
                   int max_steps) {
  float remaining = duration;
  float t = t0;
  int n = 0;
  while (remaining > 0.0f) {
    if (times) {
      if (n >= max_steps) return -1;
      times[n] = t;
    }
    n++;
    t += dt;
    remaining -= dt;
  }
  return n;
}

static float nrm3f(float x, float y, float z) {
  return sqrtf(fmaf(z, z, fmaf(y, y, x * x)));
}

/* ------------------------------------------------------------ threading */
typedef struct {
  const orc_topology* topo;
  const orc_mattab* mt;
  const orc_env* env;
  float dt, t0;
  int nsteps;
  void* pos;
  void* vel;
  void* stress;
  int mode;  /* 0 xpbd32, 1 xpbd64, 2 force32, 3 force64 */
  int r_begin, r_end;
} job_t;

static void xpbd32_robot(job_t* j, int r);
static void xpbd64_robot(job_t* j, int r);
static void force32_robot(job_t* j, int r);
static void force64_robot(job_t* j, int r);

static void* worker(void* arg) {
  job_t* j = (job_t*)arg;
  for (int r = j->r_begin; r < j->r_end; r++) {
    switch (j->mode) {
      case 0: xpbd32_robot(j, r); break;
      case 1: xpbd64_robot(j, r); break;
      case 2: force32_robot(j, r); break;
      default: force64_robot(j, r); break;
    }
  }
  return NULL;
}

static int run_jobs(job_t proto, int nthreads) {
  int nr = proto.topo->n_robots;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > nr) nthreads = nr > 0 ? nr : 1;
  job_t* jobs = (job_t*)malloc(sizeof(job_t) * (size_t)nthreads);
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
  /* balance by spring count */
  const uint32_t* so = proto.topo->spring_off;
  double total = (double)so[nr] + nr;
  int r = 0;
  for (int t = 0; t < nthreads; t++) {
    jobs[t] = proto;
    jobs[t].r_begin = r;
    double target = total * (t + 1) / nthreads;
    while (r < nr && ((double)so[r + 1] + r + 1) <= target + 1e-9) r++;
    if (t == nthreads - 1) r = nr;
    jobs[t].r_end = r;
  }
  for (int t = 1; t < nthreads; t++) pthread_create(&th[t], NULL, worker, &jobs[t]);
  worker(&jobs[0]);
  for (int t = 1; t < nthreads; t++) pthread_join(th[t], NULL);
  free(jobs);
  free(th);
  return 0;
}

/* Canonical summation order of the FP32 steps: per mass, one sum from
 * +0.0f over its springs in ascending spring index, + the term where it is
 * the first endpoint, - the term where it is the second (first, then
 * second, for a spring from a mass to itself).  Air springs contribute
 * nothing.  acc must hold zeros on entry. */
/* force terms {f, Dx, Dy, Dz} per spring: sum = fmaf(+-f, D, sum) */
static void gather_force_terms(const orc_topology* tp, uint32_t sb, uint32_t ns, const float* term,
                               float* acc) {
  for (uint32_t s = 0; s < ns; s++) {
    if (tp->mat[sb + s] == 0) continue;
    uint32_t a = tp->pairs[2 * (sb + s)], b = tp->pairs[2 * (sb + s) + 1];
    const float* t = term + 4 * (size_t)s;
    for (int c = 0; c < 3; c++) acc[3 * a + c] = fmaf(t[0], t[1 + c], acc[3 * a + c]);
    for (int c = 0; c < 3; c++) acc[3 * b + c] = fmaf(-t[0], t[1 + c], acc[3 * b + c]);
  }
}

static void gather_terms(const orc_topology* tp, uint32_t sb, uint32_t ns, uint32_t nm,
                         const float* term, float* acc, float* accm) {
  (void)nm;
  (void)accm;
  for (uint32_t s = 0; s < ns; s++) {
    if (tp->mat[sb + s] == 0) continue;
    uint32_t a = tp->pairs[2 * (sb + s)], b = tp->pairs[2 * (sb + s) + 1];
    acc[3 * a] = acc[3 * a] + term[3 * s];
    acc[3 * a + 1] = acc[3 * a + 1] + term[3 * s + 1];
    acc[3 * a + 2] = acc[3 * a + 2] + term[3 * s + 2];
    acc[3 * b] = acc[3 * b] - term[3 * s];
    acc[3 * b + 1] = acc[3 * b + 1] - term[3 * s + 1];
    acc[3 * b + 2] = acc[3 * b + 2] - term[3 * s + 2];
  }
}

/* ------------------------------------------------------------ xpbd fp32 */
static void xpbd32_robot(job_t* j, int r) {
  const orc_topology* tp = j->topo;
  const orc_mattab* mt = j->mt;
  float dt = j->dt;
  uint32_t mb = tp->mass_off[r], nm = tp->mass_off[r + 1] - mb;
  uint32_t sb = tp->spring_off[r], ns = tp->spring_off[r + 1] - sb;
  float* x = (float*)j->pos + 3 * (size_t)mb;
  float* v = (float*)j->vel + 3 * (size_t)mb;
  float* st = j->stress ? (float*)j->stress + sb : NULL;
  float* xt = (float*)malloc(sizeof(float) * 3 * nm);
  float* acc = (float*)malloc(sizeof(float) * 3 * nm);
  float* term = (float*)malloc(sizeof(float) * 3 * (ns + 1));
  float* accm = (float*)malloc(sizeof(float) * 3 * nm);
  float* K = (float*)malloc(sizeof(float) * (size_t)mt->n);
  float* R = (float*)malloc(sizeof(float) * (size_t)mt->n);
  for (int m = 0; m < mt->n; m++) {
    float alpha = 1.0f / mt->k[m];
    alpha = alpha / dt;
    alpha = alpha / dt;
    K[m] = 2.0f + alpha;
  }
  float t = j->t0;
  for (int step = 0; step < j->nsteps; step++) {
    for (uint32_t i = 0; i < 3 * nm; i++) {
      float dv = v[i] * dt;
      xt[i] = x[i] + dv;
      acc[i] = 0.0f;
    }
    for (int m = 0; m < mt->n; m++) {
      float arg = fmaf(mt->omega[m], t, mt->phi[m]);
      R[m] = mt->dL0[m] * orc_sinf(arg);
    }
    for (uint32_t s = 0; s < ns; s++) {
      uint8_t m = tp->mat[sb + s];
      if (m == 0) continue;
      uint32_t a = tp->pairs[2 * (sb + s)], b = tp->pairs[2 * (sb + s) + 1];
      float lb = tp->lbar[sb + s];
      float L = fmaf(lb, R[m], lb);
      float dx = xt[3 * a] - xt[3 * b];
      float dy = xt[3 * a + 1] - xt[3 * b + 1];
      float dz = xt[3 * a + 2] - xt[3 * b + 2];
      float d = nrm3f(dx, dy, dz);
      float den = d + ORC_EPS;
      float nx = dx / den, ny = dy / den, nz = dz / den;
      float C = d - L;
      float lambda = -C / K[m];
      float px = lambda * nx, py = lambda * ny, pz = lambda * nz;
      if (st) st[s] = st[s] + lambda / lb;
      term[3 * s] = px; term[3 * s + 1] = py; term[3 * s + 2] = pz;
    }
    gather_terms(tp, sb, ns, nm, term, acc, accm);
    for (uint32_t i = 0; i < 3 * nm; i++) {
      float xn = xt[i] + acc[i];
      float diff = xn - x[i];
      v[i] = (float)((0.99 * (double)diff) / (double)dt);
      x[i] = xn;
    }
    t = t + dt;
  }
  free(xt); free(acc); free(term); free(accm); free(K); free(R);
}

/* ------------------------------------------------------------ xpbd fp64 */
static void xpbd64_robot(job_t* j, int r) {
  const orc_topology* tp = j->topo;
  const orc_mattab* mt = j->mt;
  double dt = (double)j->dt;
  uint32_t mb = tp->mass_off[r], nm = tp->mass_off[r + 1] - mb;
  uint32_t sb = tp->spring_off[r], ns = tp->spring_off[r + 1] - sb;
  double* x = (double*)j->pos + 3 * (size_t)mb;
  double* v = (double*)j->vel + 3 * (size_t)mb;
  double* st = j->stress ? (double*)j->stress + sb : NULL;
  double* xt = (double*)malloc(sizeof(double) * 3 * nm);
  double* acc = (double*)malloc(sizeof(double) * 3 * nm);
  double* K = (double*)malloc(sizeof(double) * (size_t)mt->n);
  double* R = (double*)malloc(sizeof(double) * (size_t)mt->n);
  for (int m = 0; m < mt->n; m++) K[m] = 2.0 + ((1.0 / (double)mt->k[m]) / dt) / dt;
  float t = j->t0;
  for (int step = 0; step < j->nsteps; step++) {
    for (uint32_t i = 0; i < 3 * nm; i++) {
      xt[i] = x[i] + v[i] * dt;
      acc[i] = 0.0;
    }
    for (int m = 0; m < mt->n; m++)
      R[m] = (double)mt->dL0[m] *
             sin((double)mt->omega[m] * (double)t + (double)mt->phi[m]);
    for (uint32_t s = 0; s < ns; s++) {
      uint8_t m = tp->mat[sb + s];
      if (m == 0) continue;
      uint32_t a = tp->pairs[2 * (sb + s)], b = tp->pairs[2 * (sb + s) + 1];
      double lb = (double)tp->lbar[sb + s];
      double L = lb * (1.0 + R[m]);
      double dx = xt[3 * a] - xt[3 * b];
      double dy = xt[3 * a + 1] - xt[3 * b + 1];
      double dz = xt[3 * a + 2] - xt[3 * b + 2];
      double d = sqrt(dx * dx + dy * dy + dz * dz);
      double den = d + (double)ORC_EPS;
      double lambda = -(d - L) / K[m];
      double px = lambda * dx / den, py = lambda * dy / den,
             pz = lambda * dz / den;
      if (st) st[s] += lambda / lb;
      acc[3 * a] += px; acc[3 * a + 1] += py; acc[3 * a + 2] += pz;
      acc[3 * b] -= px; acc[3 * b + 1] -= py; acc[3 * b + 2] -= pz;
    }
    for (uint32_t i = 0; i < 3 * nm; i++) {
      double xn = xt[i] + acc[i];
      v[i] = 0.99 * (xn - x[i]) / dt;
      x[i] = xn;
    }
    t = t + j->dt;
  }
  free(xt); free(acc); free(K); free(R);
}

/* ----------------------------------------------------------- force fp32 */
static void force32_robot(job_t* j, int r) {
  const orc_topology* tp = j->topo;
  const orc_mattab* mt = j->mt;
  const orc_env* e = j->env;
  float dt = e->dt;
  uint32_t mb = tp->mass_off[r], nm = tp->mass_off[r + 1] - mb;
  uint32_t sb = tp->spring_off[r], ns = tp->spring_off[r + 1] - sb;
  float* x = (float*)j->pos + 3 * (size_t)mb;
  float* v = (float*)j->vel + 3 * (size_t)mb;
  const float* ms = tp->mass + mb;
  float* acc = (float*)malloc(sizeof(float) * 3 * nm);
  float* term = (float*)malloc(sizeof(float) * 4 * (ns + 1));
  float* R = (float*)malloc(sizeof(float) * (size_t)mt->n);
  float t = j->t0;
  for (int step = 0; step < j->nsteps; step++) {
    for (uint32_t i = 0; i < 3 * nm; i++) acc[i] = 0.0f;
    for (int m = 0; m < mt->n; m++) {
      float arg = fmaf(mt->omega[m], t, mt->phi[m]);
      R[m] = mt->dL0[m] * orc_sinf(arg);
    }
    for (uint32_t s = 0; s < ns; s++) {
      uint8_t m = tp->mat[sb + s];
      if (m == 0) continue;
      uint32_t a = tp->pairs[2 * (sb + s)], b = tp->pairs[2 * (sb + s) + 1];
      float lb = tp->lbar[sb + s];
      float L = fmaf(lb, R[m], lb);
      float dx = x[3 * a] - x[3 * b];
      float dy = x[3 * a + 1] - x[3 * b + 1];
      float dz = x[3 * a + 2] - x[3 * b + 2];
      float d = nrm3f(dx, dy, dz);
      float inv = 1.0f / (d + ORC_EPS);
      float f = (mt->k[m] * (L - d)) * inv;
      term[4 * s] = f; term[4 * s + 1] = dx; term[4 * s + 2] = dy; term[4 * s + 3] = dz;
    }
    gather_force_terms(tp, sb, ns, term, acc);
    for (uint32_t i = 0; i < nm; i++) {
      float m = ms[i];
      float fx = acc[3 * i], fy = acc[3 * i + 1], fz = acc[3 * i + 2];
      float vx = v[3 * i], vy = v[3 * i + 1], vz = v[3 * i + 2];
      float mg = m * e->gravity;
      fy = fy - mg;
      if (e->drag != 0.0f) {
        fx = fx - e->drag * vx;
        fy = fy - e->drag * vy;
        fz = fz - e->drag * vz;
      }
      float y = x[3 * i + 1];
      if (e->ground_k != 0.0f && y < 0.0f) {
        float fn = e->ground_k * (-y);
        fy = fy + fn;
        float fh = sqrtf(fx * fx + fz * fz);
        if (fh <= e->mu_s * fn) {
          fx = 0.0f; fz = 0.0f; vx = 0.0f; vz = 0.0f;
        } else {
          float sc = 1.0f - (e->mu_k * fn) / fh;
          fx = fx * sc;
          fz = fz * sc;
        }
      }
      float inv = 1.0f / m;
      float ux = vx + (fx * inv) * dt;
      float uy = vy + (fy * inv) * dt;
      float uz = vz + (fz * inv) * dt;
      x[3 * i] = x[3 * i] + ux * dt;
      x[3 * i + 1] = x[3 * i + 1] + uy * dt;
      x[3 * i + 2] = x[3 * i + 2] + uz * dt;
      v[3 * i] = e->damping * ux;
      v[3 * i + 1] = e->damping * uy;
      v[3 * i + 2] = e->damping * uz;
    }
    t = t + dt;
  }
  free(acc); free(term); free(R);
}

/* ----------------------------------------------------------- force fp64 */
static void force64_robot(job_t* j, int r) {
  const orc_topology* tp = j->topo;
  const orc_mattab* mt = j->mt;
  const orc_env* e = j->env;
  double dt = (double)e->dt;
  uint32_t mb = tp->mass_off[r], nm = tp->mass_off[r + 1] - mb;
  uint32_t sb = tp->spring_off[r], ns = tp->spring_off[r + 1] - sb;
  double* x = (double*)j->pos + 3 * (size_t)mb;
  double* v = (double*)j->vel + 3 * (size_t)mb;
  const float* ms = tp->mass + mb;
  double* acc = (double*)malloc(sizeof(double) * 3 * nm);
  double* R = (double*)malloc(sizeof(double) * (size_t)mt->n);
  double g = e->gravity, kg = e->ground_k, mus = e->mu_s, muk = e->mu_k,
         c = e->drag, damp = e->damping;
  float t = j->t0;
  for (int step = 0; step < j->nsteps; step++) {
    for (uint32_t i = 0; i < 3 * nm; i++) acc[i] = 0.0;
    for (int m = 0; m < mt->n; m++)
      R[m] = (double)mt->dL0[m] *
             sin((double)mt->omega[m] * (double)t + (double)mt->phi[m]);
    for (uint32_t s = 0; s < ns; s++) {
      uint8_t m = tp->mat[sb + s];
      if (m == 0) continue;
      uint32_t a = tp->pairs[2 * (sb + s)], b = tp->pairs[2 * (sb + s) + 1];
      double lb = (double)tp->lbar[sb + s];
      double L = lb * (1.0 + R[m]);
      double dx = x[3 * a] - x[3 * b];
      double dy = x[3 * a + 1] - x[3 * b + 1];
      double dz = x[3 * a + 2] - x[3 * b + 2];
      double d = sqrt(dx * dx + dy * dy + dz * dz);
      double den = d + (double)ORC_EPS;
      double f = (double)mt->k[m] * (L - d) / den;
      acc[3 * a] += f * dx; acc[3 * a + 1] += f * dy; acc[3 * a + 2] += f * dz;
      acc[3 * b] -= f * dx; acc[3 * b + 1] -= f * dy; acc[3 * b + 2] -= f * dz;
    }
    for (uint32_t i = 0; i < nm; i++) {
      double m = ms[i];
      double fx = acc[3 * i], fy = acc[3 * i + 1] - m * g, fz = acc[3 * i + 2];
      double vx = v[3 * i], vy = v[3 * i + 1], vz = v[3 * i + 2];
      if (c != 0.0) { fx -= c * vx; fy -= c * vy; fz -= c * vz; }
      double y = x[3 * i + 1];
      if (kg != 0.0 && y < 0.0) {
        double fn = kg * (-y);
        fy += fn;
        double fh = sqrt(fx * fx + fz * fz);
        if (fh <= mus * fn) {
          fx = fz = 0.0; vx = vz = 0.0;
        } else {
          double sc = 1.0 - muk * fn / fh;
          fx *= sc; fz *= sc;
        }
      }
      double ux = vx + fx / m * dt, uy = vy + fy / m * dt, uz = vz + fz / m * dt;
      x[3 * i] += ux * dt; x[3 * i + 1] += uy * dt; x[3 * i + 2] += uz * dt;
      v[3 * i] = damp * ux; v[3 * i + 1] = damp * uy; v[3 * i + 2] = damp * uz;
    }
    t = t + e->dt;
  }
  free(acc); free(R);
}

/* ----------------------------------------------------------- entry points */
int orc_xpbd_f32(const orc_topology* topo, const orc_mattab* mt, float dt,
                 float t0, int nsteps, float* pos, float* vel, float* stress,
                 int nthreads) {
  job_t j;
  memset(&j, 0, sizeof j);
  j.topo = topo; j.mt = mt; j.dt = dt; j.t0 = t0; j.nsteps = nsteps;
  j.pos = pos; j.vel = vel; j.stress = stress; j.mode = 0;
  return run_jobs(j, nthreads);
}

int orc_xpbd_f64(const orc_topology* topo, const orc_mattab* mt, float dt,
                 float t0, int nsteps, double* pos, double* vel,
                 double* stress, int nthreads) {
  job_t j;
  memset(&j, 0, sizeof j);
  j.topo = topo; j.mt = mt; j.dt = dt; j.t0 = t0; j.nsteps = nsteps;
  j.pos = pos; j.vel = vel; j.stress = stress; j.mode = 1;
  return run_jobs(j, nthreads);
}

int orc_force_f32(const orc_topology* topo, const orc_mattab* mt,
                  const orc_env* env, float t0, int nsteps, float* pos,
                  float* vel, int nthreads) {
  job_t j;
  memset(&j, 0, sizeof j);
  j.topo = topo; j.mt = mt; j.env = env; j.dt = env->dt; j.t0 = t0;
  j.nsteps = nsteps; j.pos = pos; j.vel = vel; j.mode = 2;
  return run_jobs(j, nthreads);
}

int orc_force_f64(const orc_topology* topo, const orc_mattab* mt,
                  const orc_env* env, float t0, int nsteps, double* pos,
                  double* vel, int nthreads) {
  job_t j;
  memset(&j, 0, sizeof j);
  j.topo = topo; j.mt = mt; j.env = env; j.dt = env->dt; j.t0 = t0;
  j.nsteps = nsteps; j.pos = pos; j.vel = vel; j.mode = 3;
  return run_jobs(j, nthreads);
}
