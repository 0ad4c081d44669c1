/*
 * nbody_amd.h - C ABI of the MI355X (gfx950) all-pairs gravity step.
 *
 * This is the drop-in boundary for the reference's simulation layer:
 *   - class simulation::DiskGalaxySimulator  (src/simulator.cuh:129-160)
 *   - kernel particle_interaction<CalculationMethod>  (src/simulator.cu:186-229)
 *   - SimParam  (src/sim_param.hpp:16-39)
 *
 * Everything is plain C: pointers, sizes and status codes.  No HIP or torch
 * type appears in a signature; streams are passed as `void *` (a hipStream_t,
 * or NULL for the null stream).  Functions never throw and never exit: a
 * nonzero status is returned and nb_last_error() holds a message for the
 * calling thread.
 *
 * Arrays are structure-of-arrays (SoA): one float[n] per coordinate, the
 * layout of ParticleData / ParticleData_d (src/simulator.cuh:75-98).
 *
 * Two levels:
 *   nb_sim_*     an owning simulator handle: host arrays in, host arrays out,
 *                device-resident state between steps (the DiskGalaxySimulator
 *                contract).
 *   nb_*_dev     stream-ordered launches on caller-owned device buffers, for
 *                callers that manage device memory themselves (PyTorch, graphs).
 */
#ifndef NBODY_AMD_H
#define NBODY_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define NB_OK 0
#define NB_ERR_INVALID_ARG 1   /* bad size, null pointer, unknown option       */
#define NB_ERR_NO_DEVICE 2     /* no HIP device visible                         */
#define NB_ERR_HIP 3           /* a HIP runtime call failed                     */
#define NB_ERR_OUT_OF_MEMORY 4 /* device or host allocation failed              */
#define NB_ERR_UNSUPPORTED 5   /* option combination not implemented            */

/* CalculationMethod (src/sim_param.hpp:8-11).  PREDICATED reproduces the
 * mounted reference exactly: force += r*inv*(j == id) (src/simulator.cu:209),
 * which keeps only the self term. */
#define NB_CALC_BRANCH 0
#define NB_CALC_PREDICATED 1

/* Force-sum accumulation mode (extension; BRANCH only). */
#define NB_ACC_PLAIN 0 /* one fp32 FMA per term (default)                      */
#define NB_ACC_KAHAN 1 /* Kahan-compensated fp32 sum of the per-particle force
                          (target-sharded shapes, and symmetric shapes with 2
                          or 4 particles per lane)                             */

/* SimParam, field for field (src/sim_param.hpp:30-39). */
typedef struct nb_params {
  float G;                    /* gravitational parameter                    */
  float dt;                   /* time step                                  */
  size_t num_particles;       /* N                                          */
  size_t num_frames;          /* frames for a driver loop                   */
  int sim_iterations_per_frame; /* kernel launches per stepSim()            */
  float damping;              /* velocity drag factor (1 = none)            */
  float dist_eps;             /* softening added to |r|^2                   */
  int gw_size;                /* reference work-group size (a hint here)    */
  int calc_method;            /* NB_CALC_BRANCH or NB_CALC_PREDICATED       */
} nb_params;

/* SimParam::SimParam() defaults (src/sim_param.cpp:12-22). */
void nb_params_default(nb_params *p);

/* Message for the last failed call on this thread ("" if none). */
const char *nb_last_error(void);

/* Number of visible HIP devices (0 without a GPU; never fails). */
int nb_device_count(void);

/* Name of device 0 (DiskGalaxySimulator::getDeviceName, src/simulator.cu:36-45).
 * Writes "Unknown Device" when the query fails, as the reference does. */
int nb_device_name(char *buf, size_t len);

/* ---- initial conditions (host) ------------------------------------------ */
/* Disk in the draw order of DiskGalaxySimulator::randomParticlePos and
 * initialParticleVel (src/simulator.cu:131-158): per particle t then s, then
 * all z; std::mt19937(seed) with uniform_real_distribution<double>.
 * seed 5489 is the reference's default-constructed engine.
 * flags: NB_IC_TRIG_DOUBLE evaluates cos/sin in double (the default is the
 * float overload); NB_IC_LENGTH_F32 uses the fp32 length() of the SYCL twin
 * (src_sycl/simulator.dp.cpp:299-301) instead of the double pow form
 * (src/simulator.cu:170-172). */
#define NB_IC_TRIG_DOUBLE 1
#define NB_IC_LENGTH_F32 2
int nb_ic_disk(size_t n, uint32_t seed, int flags, float *px, float *py,
               float *pz, float *vx, float *vy, float *vz);

/* Disk in the draw order of gen.cpp randomParticlePos/randomParticleVel
 * (src/gen.cpp:15-34): t, s, z drawn per particle; glm-style normalize. */
int nb_ic_disk_gen(size_t n, uint32_t seed, float *px, float *py, float *pz,
                   float *vx, float *vy, float *vz);

/* Plummer sphere (no reference counterpart): scale radius a, unit masses,
 * isotropic velocities in virial equilibrium for gravitational parameter
 * g_mass = G * total mass, sampled as in Aarseth, Henon & Wielen (1974),
 * shifted to the centre-of-mass frame. */
int nb_ic_plummer(size_t n, uint32_t seed, float a, float g_mass, float *px,
                  float *py, float *pz, float *vx, float *vy, float *vz);

/* ---- owning simulator handle (DiskGalaxySimulator) ---------------------- */
typedef struct nb_sim nb_sim;

/* Copies the caller's SoA arrays (p->num_particles each) to the device.
 * Replaces the DiskGalaxySimulator(SimParam) constructor's sendToDevice
 * (src/simulator.cu:24-34, 79-103) for caller-supplied initial conditions. */
int nb_sim_create(const nb_params *p, const float *px, const float *py,
                  const float *pz, const float *vx, const float *vy,
                  const float *vz, nb_sim **out);

/* stepSim() without the readback (src/simulator.cu:47-72): launches
 * p->sim_iterations_per_frame steps, ping-pongs positions, synchronises and
 * records the wall-clock frame time. */
int nb_sim_step(nb_sim *s);

/* Launches `iters` steps without synchronising (stream-ordered). */
int nb_sim_advance(nb_sim *s, int iters);

/* Blocks until every launched step has finished. */
int nb_sim_sync(nb_sim *s);

/* recvFromDevice (src/simulator.cu:106-129): current positions / velocities. */
int nb_sim_read_pos(nb_sim *s, float *x, float *y, float *z);
int nb_sim_read_vel(nb_sim *s, float *x, float *y, float *z);

/* Per-frame readback through pinned host memory: nb_sim_stage_readback
 * enqueues one kernel that stores the current positions and velocities into
 * a pinned, device-mapped buffer owned by the handle and returns at once;
 * nb_sim_staged waits for it and returns six pointers (x, y, z, vx, vy, vz)
 * valid until the next stage call or nb_sim_destroy.  The reference issues
 * six synchronous pageable copies per frame instead
 * (src/simulator.cu:106-129). */
int nb_sim_stage_readback(nb_sim *s);
int nb_sim_staged(nb_sim *s, const float **arrays);

/* Positions (which = 0) or velocities (which = 1) as interleaved
 * vec4(x, y, z, w) into a host array of 4 n floats: the renderer's buffer
 * format (src/renderer_gl.cpp:156-172, shaders/gl/main.vert:7-8), packed on
 * the device. */
int nb_sim_read_vec4(nb_sim *s, int which, float w, float *out);

/* getLastStepTime (src/simulator.cuh:134): ms of the last nb_sim_step, from
 * before the first launch until the device synchronise. */
float nb_sim_last_step_ms(const nb_sim *s);

/* Device time of the last nb_sim_step's launches, from HIP events (ms). */
float nb_sim_last_kernel_ms(const nb_sim *s);

size_t nb_sim_num_particles(const nb_sim *s); /* getNumParticles */
int nb_sim_calc_method(const nb_sim *s);      /* getCM             */
int nb_sim_gw_size(const nb_sim *s);          /* getGwSize         */

/* Selects the accumulation mode (NB_ACC_*) used by later steps. */
int nb_sim_set_accumulation(nb_sim *s, int mode);

/* Changes simIterationsPerFrame for later nb_sim_step calls. */
int nb_sim_set_frame_iterations(nb_sim *s, int iters);

/* Frames of two or more steps (nb_sim_step, nb_sim_advance) are captured once
 * as a hipGraph per (steps, launch shape, accumulation, start buffer) and
 * replayed; enable = 0 launches every step directly (results are the same
 * bits).  No counterpart in the reference, whose stepSim launches each step
 * (src/simulator.cu:59-66).  Default: on. */
int nb_sim_set_graph(nb_sim *s, int enable);

/* Selects a forward launch shape (see nb_forward_config_*) for later steps;
 * -1 restores the library's choice. */
int nb_sim_set_config(nb_sim *s, int cfg);
int nb_sim_config(const nb_sim *s);

int nb_sim_destroy(nb_sim *s);

/* ---- stream-ordered device entry points --------------------------------- */
/* One launch of particle_interaction (src/simulator.cu:186-229):
 *   reads pos (px,py,pz), updates vel (vx,vy,vz) in place, writes the next
 *   positions (nx,ny,nz).  pos and next must not overlap.  n = particle count.
 * cfg selects a tuned launch shape (-1 = the library's choice for n);
 * acc is NB_ACC_*. */
int nb_forward_step_dev(const nb_params *p, const float *px, const float *py,
                        const float *pz, float *vx, float *vy, float *vz,
                        float *nx, float *ny, float *nz, size_t n, int cfg,
                        int acc, void *stream);

/* nb_forward_step_dev without any allocation (no counterpart in the
 * reference: it is for callers that cannot allocate device memory, such as
 * code under graph capture or a tight serving loop).  The same arguments,
 * shapes and results as nb_forward_step_dev, bit for bit; the partial sums
 * of shapes that keep them live in a static 64 MiB device buffer of the
 * library (one per GPU, held until the process ends), which covers the
 * default shapes of plain sums up to n = 26624 (the reference's default
 * 12800 and the 25600 of its CPU configuration among them; Kahan sums up to
 * n = 18816).  Nothing in it needs clearing between calls: every partial
 * sum is written before it is read.  A shape whose partial sums
 * exceed it is refused with NB_ERR_UNSUPPORTED.  Every call on one GPU uses
 * that GPU's buffer: calls that overlap on one GPU (different streams or
 * host threads) race and must be ordered by the caller. */
int nb_forward_step_noalloc_dev(const nb_params *p, const float *px,
                                const float *py, const float *pz, float *vx,
                                float *vy, float *vz, float *nx, float *ny,
                                float *nz, size_t n, int cfg, int acc,
                                void *stream);

/* Number of compiled forward launch shapes, and a description of one:
 * block threads, targets per thread, LDS tile length, j-split count.
 * For a symmetric shape: 256 threads, particles per lane, super-tile length,
 * super-tiles per work item. */
int nb_forward_config_count(void);
int nb_forward_config_info(int cfg, int *block, int *unroll, int *tile,
                           int *split);
/* The shape the library picks (cfg = -1) for n particles with these
 * parameters and accumulation mode; -1 (and nb_last_error) for bad
 * arguments. */
int nb_forward_default_config(const nb_params *p, size_t n, int acc);

/* Decomposition of a forward launch shape (same step, same per-pair
 * arithmetic; no counterpart in the reference, whose kernel is the
 * target-sharded one):
 *   NB_SCHEME_TARGETS    each thread sums all sources for its targets;
 *   NB_SCHEME_SYMMETRIC  each pair term is evaluated once and applied to both
 *                        particles (Newton's third law); BRANCH (plain or
 *                        Kahan sums), n <= 2^20 (2^18 with 2 per lane; 2^24
 *                        with 4 super-tiles per item, in bands of source
 *                        super-tiles past 2^20, past 2^18 with 2 per lane,
 *                        past 2^21 with 8 per lane and plain sums),
 *                        scratch growing as n^2 (0.4 GB at n = 2^18 with 8
 *                        per lane, 16 GB at 2^21, at most 8 GB per band,
 *                        twice that with Kahan sums);
 *                        work item = a pair of 4-tile super-tiles;
 *   NB_SCHEME_SYMMETRIC_TILES  the same with one tile pair per wave, for
 *                        small n (n <= 2^17).
 * Returns the scheme, or -1 (and sets nb_last_error) for a bad index. */
#define NB_SCHEME_TARGETS 0
#define NB_SCHEME_SYMMETRIC 1
#define NB_SCHEME_SYMMETRIC_TILES 2
int nb_forward_config_scheme(int cfg);

/* Scratch of the nb_*_dev entry points (no reference counterpart: the
 * reference kernel keeps no partial sums).  Shapes with partial sums
 * (source-split target-sharded shapes, the symmetric shapes, the fp64
 * diagnostics) use a buffer the library allocates on first use, one per
 * (device, stream): calls on one stream reuse it, calls on different streams
 * never share it, and a call that needs more grows it (after waiting for the
 * stream).  A call that would grow it while its stream is being captured
 * into a graph fails with NB_ERR_UNSUPPORTED; make the call once before
 * capturing.
 *   nb_scratch_release  waits for all work on the current device, then frees
 *                       every such buffer of the device (they are allocated
 *                       again on demand);
 *   nb_scratch_held     bytes the library holds as such buffers on the
 *                       current device. */
int nb_scratch_release(void);
int nb_scratch_held(size_t *bytes);


/* ---- adjoint (extension; no reference counterpart) ---------------------- */
/* Three device arrays of one SoA vector field. */
typedef struct nb_vec3 {
  float *x;
  float *y;
  float *z;
} nb_vec3;

/* Reverse-mode derivative of one step (x, v) -> (x', v') of
 * particle_interaction's integrator (src/simulator.cu:214-228):
 *   g = lam_v_next + dt * lam_x_next,  lam_v = damping * g,
 *   lam_x = lam_x_next + dt * G * sum_{j != k} M_kj (g_j - g_k),
 *   M = K (I - 3 r r^T / d),  d = |r|^2 + eps,  K = d^-3/2,
 * evaluated at positions pos (x_n) in one pass over all pairs.  BRANCH only.
 * lam_x / lam_v must not alias the inputs.  cfg: -1 = library's shape. */
int nb_adjoint_step_dev(const nb_params *p, const nb_vec3 *pos,
                        const nb_vec3 *lam_x_next, const nb_vec3 *lam_v_next,
                        const nb_vec3 *lam_x, const nb_vec3 *lam_v, size_t n,
                        int cfg, void *stream);

/* The forward step and the adjoint step at the same positions in one
 * launch: vel is updated in place and pos_next written (as
 * nb_forward_step_dev), and the cotangents propagated (as
 * nb_adjoint_step_dev).  The pair quantities are computed once. */
int nb_fwd_adjoint_step_dev(const nb_params *p, const nb_vec3 *pos,
                            const nb_vec3 *vel, const nb_vec3 *pos_next,
                            const nb_vec3 *lam_x_next,
                            const nb_vec3 *lam_v_next, const nb_vec3 *lam_x,
                            const nb_vec3 *lam_v, size_t n, int cfg,
                            void *stream);

int nb_adjoint_config_count(void);
int nb_adjoint_config_info(int cfg, int *block, int *unroll, int *tile,
                           int *split);
/* The adjoint shape the library picks (cfg = -1) for n particles with these
 * parameters; -1 (and nb_last_error) for a NULL p. */
int nb_adjoint_default_config(const nb_params *p, size_t n);
/* Scheme of an adjoint launch shape (NB_SCHEME_*; the symmetric shapes take
 * n <= 2^20 (2^18 with 2 particles per lane), 2^24 with 4 super-tiles per
 * item (in bands of source super-tiles past those limits, 16 GB of partials
 * per band for the fused step), 2^17 for tile pairs), or -1 for a bad
 * index. */
int nb_adjoint_config_scheme(int cfg);

/* ---- differentiable trajectory with checkpoints (extension) ------------- */
/* States are [6][n] float blocks: x, y, z, vx, vy, vz (n floats each).
 * The forward pass runs `steps` steps from state0, stores the state every
 * `every` steps (checkpoint slot s = state at step s * every) and writes
 * the final state.  The backward pass takes the cotangent of the final state
 * ([6][n]: lam_x then lam_v), recomputes each segment's positions from its
 * checkpoint and runs the adjoint steps in reverse, writing the cotangent of
 * state0.  Both are stream-ordered and allocate nothing: the caller passes
 * checkpoint storage and a workspace of the sizes below (in floats). */
size_t nb_traj_checkpoint_floats(size_t n, int steps, int every);
size_t nb_traj_workspace_floats(size_t n, int every);
int nb_traj_forward_dev(const nb_params *p, const float *state0,
                        float *state_final, int steps, int every,
                        float *checkpoints, float *workspace, size_t n,
                        void *stream);
int nb_traj_backward_dev(const nb_params *p, const float *checkpoints,
                         const float *lam_final, float *lam0, int steps,
                         int every, float *workspace, size_t n, void *stream);

/* ---- batched independent systems (extension) ---------------------------- */
/* Many independent systems advanced together.  System s has
 * params[s].num_particles particles (BRANCH; its own G, dt, damping,
 * dist_eps); the host arrays hold all systems back to back.  One step is
 * one launch of a persistent kernel that drains a device work queue of
 * (system, target block) items, longest first; steps are replayed from a
 * captured two-step hipGraph. */
typedef struct nb_batch nb_batch;
int nb_batch_create(int num_systems, const nb_params *params, const float *px,
                    const float *py, const float *pz, const float *vx,
                    const float *vy, const float *vz, nb_batch **out);
/* `steps` steps through the captured graph; synchronises and records the
 * device time in nb_batch_last_ms. */
int nb_batch_step(nb_batch *b, int steps);
/* The same steps as direct launches (no graph). */
int nb_batch_step_direct(nb_batch *b, int steps);
/* Baseline: one nb_forward_step-style launch per system per step. */
int nb_batch_step_separate(nb_batch *b, int steps);
/* Times every compiled launch shape on this batch (trial_steps each, the
 * state restored after each trial) and keeps the fastest. */
int nb_batch_autotune(nb_batch *b, int trial_steps, int *best_cfg);
int nb_batch_set_config(nb_batch *b, int cfg);
int nb_batch_config(const nb_batch *b);
int nb_batch_config_count(void);
int nb_batch_config_info(int cfg, int *block, int *unroll, int *tile);
/* Scheme of a batch shape (NB_SCHEME_TARGETS: work item = a block of one
 * system's targets; NB_SCHEME_SYMMETRIC: work item = a super-tile pair of one
 * system, each pair term evaluated once), or -1 for a bad index. */
int nb_batch_config_scheme(int cfg);
float nb_batch_last_ms(const nb_batch *b);
size_t nb_batch_num_particles(const nb_batch *b);
int nb_batch_read_pos(nb_batch *b, float *x, float *y, float *z);
int nb_batch_read_vel(nb_batch *b, float *x, float *y, float *z);
int nb_batch_destroy(nb_batch *b);

/* ---- render hand-off ----------------------------------------------------- */
/* SoA -> AoS pack: out[4i..4i+3] = (x_i, y_i, z_i, w) for i < n.  out is a
 * device array of 4 n floats, 16 B aligned; soa and out must not overlap. */
int nb_pack_vec4_dev(const nb_vec3 *soa, float w, float *out, size_t n,
                     void *stream);

/* ---- fp64 diagnostics (extension) --------------------------------------- */
/* fp64 oracle force (BRANCH semantics, the step's eps) on the targets
 * idx[0..m) (device int32 indices; an index outside [0, n) yields NaN) from
 * all n float sources: F_i = sum_{j != i} r (|r|^2 + eps)^{-3/2}, r formed in
 * double.  fx, fy, fz: device double arrays of m values. */
int nb_forces_f64_dev(const nb_params *p, const nb_vec3 *pos,
                      const int32_t *idx, size_t m, double *fx, double *fy,
                      double *fz, size_t n, void *stream);

/* fp64 totals into the device array out[5]: kinetic sum |v|^2 / 2,
 * potential -G sum_{i<j} (|r|^2 + eps)^{-1/2} (whose negative gradient is
 * the step's force), momentum sum v (unit masses).  Deterministic order. */
int nb_energy_momentum_dev(const nb_params *p, const nb_vec3 *pos,
                           const nb_vec3 *vel, size_t n, double *out,
                           void *stream);

#ifdef __cplusplus
}
#endif

#endif /* NBODY_AMD_H */
