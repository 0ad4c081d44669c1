// C ABI of the adjoint extension: single adjoint / fused steps and the
// checkpointed differentiable trajectory (host orchestration of the forward
// and adjoint kernels; all work is stream-ordered on the caller's stream and
// nothing is allocated here).

#include <algorithm>

#include "nb_internal.hpp"

using namespace nb;

namespace {

constexpr size_t kMaxN = (size_t)1 << 28;

StepConsts consts(const nb_params *p) {
  StepConsts c;
  c.eps = p->dist_eps;
  c.damping = p->damping;
  c.dt = p->dt;
  c.G = p->G;
  return c;
}

int check_common(const nb_params *p, size_t n) {
  if (!p) return set_error(NB_ERR_INVALID_ARG, "params is NULL");
  if (p->calc_method != NB_CALC_BRANCH)
    return set_error(NB_ERR_UNSUPPORTED,
                     "the adjoint is defined for CalculationMethod BRANCH");
  if (n > kMaxN) return set_error(NB_ERR_INVALID_ARG, "n exceeds 2^28");
  return NB_OK;
}

bool vec_ok(const nb_vec3 *v) { return v && v->x && v->y && v->z; }

bool overlaps(const nb_vec3 *a, const nb_vec3 *b) {
  const float *pa[3] = {a->x, a->y, a->z};
  const float *pb[3] = {b->x, b->y, b->z};
  for (const float *u : pa)
    for (const float *w : pb)
      if (u == w) return true;
  return false;
}

int adjoint_common(const nb_params *p, const nb_vec3 *pos, const nb_vec3 *vel,
                   const nb_vec3 *nxt, const nb_vec3 *lxn, const nb_vec3 *lvn,
                   const nb_vec3 *lx, const nb_vec3 *lv, size_t n, int cfg,
                   bool fused, void *stream) {
  int st = check_common(p, n);
  if (st) return st;
  if (cfg < -1 || cfg >= adjoint_config_count())
    return set_error(NB_ERR_INVALID_ARG, "adjoint config out of range");
  if (n == 0) return NB_OK;
  if (!vec_ok(pos) || !vec_ok(lxn) || !vec_ok(lvn) || !vec_ok(lx) ||
      !vec_ok(lv) || (fused && (!vec_ok(vel) || !vec_ok(nxt))))
    return set_error(NB_ERR_INVALID_ARG, "device pointer is NULL");
  if (overlaps(lx, lxn) || overlaps(lx, lvn) || overlaps(lv, lxn) ||
      overlaps(lv, lvn) || overlaps(lx, lv) ||
      (fused && (overlaps(nxt, pos) || overlaps(vel, pos))))
    return set_error(NB_ERR_INVALID_ARG, "outputs must not alias inputs");
  if (cfg < 0) cfg = adjoint_default_config(n, !self_term_vanishes(p->dist_eps));
  const hipStream_t s = static_cast<hipStream_t>(stream);
  ScratchLease scratch; // held until the launches are enqueued
  if ((st = acquire_scratch(adjoint_scratch_bytes(cfg, n), s, scratch))) return st;
  const float *P[3] = {pos->x, pos->y, pos->z};
  const float *LXN[3] = {lxn->x, lxn->y, lxn->z};
  const float *LVN[3] = {lvn->x, lvn->y, lvn->z};
  float *LX[3] = {lx->x, lx->y, lx->z};
  float *LV[3] = {lv->x, lv->y, lv->z};
  float *V[3] = {fused ? vel->x : nullptr, fused ? vel->y : nullptr,
                 fused ? vel->z : nullptr};
  float *NX[3] = {fused ? nxt->x : nullptr, fused ? nxt->y : nullptr,
                  fused ? nxt->z : nullptr};
  return launch_adjoint(consts(p), !self_term_vanishes(p->dist_eps), fused,
                        cfg, P, LXN, LVN, LX, LV, V, NX, (int)n, scratch.get(), s);
}

// Workspace layout (floats) of the trajectory passes.
struct TrajLayout {
  size_t pos_a, pos_b, vel, replay, lam_a, lam_b, scratch, total;
};

TrajLayout traj_layout(size_t n, int every) {
  TrajLayout L;
  const size_t n3 = 3 * n;
  L.pos_a = 0;
  L.pos_b = n3;
  L.vel = 2 * n3;
  L.replay = 3 * n3;
  L.lam_a = L.replay + (size_t)every * n3;
  L.lam_b = L.lam_a + 2 * n3;
  L.scratch = L.lam_b + 2 * n3;
  // the forward shape depends on the pair mode: size for both candidates
  const int cf = forward_default_config(n, MODE_PLAIN);
  const int cs = forward_default_config(n, MODE_PLAIN_SKIP);
  const int ca = adjoint_default_config(n, false);
  const int ck = adjoint_default_config(n, true);
  const size_t sb = std::max({forward_scratch_bytes(cf, n, MODE_PLAIN),
                              forward_scratch_bytes(cs, n, MODE_PLAIN_SKIP),
                              adjoint_scratch_bytes(ca, n),
                              adjoint_scratch_bytes(ck, n)});
  L.total = L.scratch + (sb + sizeof(float) - 1) / sizeof(float);
  return L;
}

int copy_d2d(float *dst, const float *src, size_t floats, hipStream_t s) {
  return hip_check(hipMemcpyAsync(dst, src, floats * sizeof(float),
                                  hipMemcpyDeviceToDevice, s),
                   "device-to-device copy");
}

int forward_on(const nb_params *p, float *from, float *to, float *vel,
               size_t n, float *scratch, hipStream_t s) {
  const int mode = pair_mode(p->calc_method, NB_ACC_PLAIN, p->dist_eps);
  const int cfg = forward_default_config(n, mode);
  return launch_forward(consts(p), mode, cfg, from, from + n, from + 2 * n,
                        vel, vel + n, vel + 2 * n, to, to + n, to + 2 * n,
                        (int)n, scratch, s);
}

int traj_args(const nb_params *p, int steps, int every, size_t n) {
  int st = check_common(p, n);
  if (st) return st;
  if (steps < 0) return set_error(NB_ERR_INVALID_ARG, "steps < 0");
  if (every < 1) return set_error(NB_ERR_INVALID_ARG, "every < 1");
  return NB_OK;
}

} // namespace

extern "C" {

int nb_adjoint_step_dev(const nb_params *p, const nb_vec3 *pos,
                        const nb_vec3 *lam_x_next, const nb_vec3 *lam_v_next,
                        const nb_vec3 *lam_x, const nb_vec3 *lam_v, size_t n,
                        int cfg, void *stream) {
  return adjoint_common(p, pos, nullptr, nullptr, lam_x_next, lam_v_next,
                        lam_x, lam_v, n, cfg, false, stream);
}

int nb_fwd_adjoint_step_dev(const nb_params *p, const nb_vec3 *pos,
                            const nb_vec3 *vel, const nb_vec3 *pos_next,
                            const nb_vec3 *lam_x_next,
                            const nb_vec3 *lam_v_next, const nb_vec3 *lam_x,
                            const nb_vec3 *lam_v, size_t n, int cfg,
                            void *stream) {
  return adjoint_common(p, pos, vel, pos_next, lam_x_next, lam_v_next, lam_x,
                        lam_v, n, cfg, true, stream);
}

int nb_adjoint_config_count(void) { return adjoint_config_count(); }

int nb_adjoint_config_info(int cfg, int *block, int *unroll, int *tile,
                           int *split) {
  if (cfg < 0 || cfg >= adjoint_config_count())
    return set_error(NB_ERR_INVALID_ARG, "adjoint config out of range");
  const FwdConfig &c = adjoint_config(cfg);
  if (block) *block = c.block;
  if (unroll) *unroll = c.unroll;
  if (tile) *tile = c.tile;
  if (split) *split = c.split;
  return NB_OK;
}

int nb_adjoint_config_scheme(int cfg) {
  if (cfg < 0 || cfg >= adjoint_config_count()) {
    set_error(NB_ERR_INVALID_ARG, "adjoint config out of range");
    return -1;
  }
  return adjoint_config(cfg).scheme;
}

int nb_adjoint_default_config(const nb_params *p, size_t n) {
  if (!p) {
    set_error(NB_ERR_INVALID_ARG, "params is NULL");
    return -1;
  }
  return adjoint_default_config(n, !self_term_vanishes(p->dist_eps));
}

size_t nb_traj_checkpoint_floats(size_t n, int steps, int every) {
  if (steps <= 0 || every < 1) return 0;
  const size_t segs = ((size_t)steps + every - 1) / every;
  return segs * 6 * n;
}

size_t nb_traj_workspace_floats(size_t n, int every) {
  if (every < 1) return 0;
  return traj_layout(n, every).total;
}

int nb_traj_forward_dev(const nb_params *p, const float *state0,
                        float *state_final, int steps, int every,
                        float *checkpoints, float *workspace, size_t n,
                        void *stream) {
  int st = traj_args(p, steps, every, n);
  if (st || n == 0) return st;
  if (!state0 || !state_final || !workspace || (steps > 0 && !checkpoints))
    return set_error(NB_ERR_INVALID_ARG, "device pointer is NULL");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t n3 = 3 * n;
  const TrajLayout L = traj_layout(n, every);
  float *pos[2] = {workspace + L.pos_a, workspace + L.pos_b};
  float *vel = workspace + L.vel;
  float *scratch = workspace + L.scratch;
  if ((st = copy_d2d(pos[0], state0, n3, s)) ||
      (st = copy_d2d(vel, state0 + n3, n3, s)))
    return st;
  int cur = 0;
  for (int t = 0; t < steps; ++t) {
    if (t % every == 0) {
      float *slot = checkpoints + (size_t)(t / every) * 6 * n;
      if ((st = copy_d2d(slot, pos[cur], n3, s)) ||
          (st = copy_d2d(slot + n3, vel, n3, s)))
        return st;
    }
    if ((st = forward_on(p, pos[cur], pos[cur ^ 1], vel, n, scratch, s)))
      return st;
    cur ^= 1;
  }
  if ((st = copy_d2d(state_final, pos[cur], n3, s)) ||
      (st = copy_d2d(state_final + n3, vel, n3, s)))
    return st;
  return NB_OK;
}

int nb_traj_backward_dev(const nb_params *p, const float *checkpoints,
                         const float *lam_final, float *lam0, int steps,
                         int every, float *workspace, size_t n,
                         void *stream) {
  int st = traj_args(p, steps, every, n);
  if (st || n == 0) return st;
  if (!lam_final || !lam0 || !workspace || (steps > 0 && !checkpoints))
    return set_error(NB_ERR_INVALID_ARG, "device pointer is NULL");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t n3 = 3 * n;
  const TrajLayout L = traj_layout(n, every);
  float *vel = workspace + L.vel;
  float *scratch = workspace + L.scratch;
  float *lam[2] = {workspace + L.lam_a, workspace + L.lam_b};
  auto replay = [&](int m) { return workspace + L.replay + (size_t)m * n3; };
  const StepConsts c = consts(p);
  const bool skip = !self_term_vanishes(p->dist_eps);
  const int cfg_a = adjoint_default_config(n, skip);

  if ((st = copy_d2d(lam[0], lam_final, 2 * n3, s))) return st;
  int cur = 0;
  const int segs = steps > 0 ? (steps + every - 1) / every : 0;
  for (int seg = segs - 1; seg >= 0; --seg) {
    const int c0 = seg * every;
    const int len = std::min(every, steps - c0);
    const float *slot = checkpoints + (size_t)seg * 6 * n;
    if ((st = copy_d2d(replay(0), slot, n3, s)) ||
        (st = copy_d2d(vel, slot + n3, n3, s)))
      return st;
    // positions x_{c0+1} .. x_{c0+len-1} from the checkpoint, with the
    // kernel and launch shape of the forward pass (bit-identical states)
    for (int m = 0; m + 1 < len; ++m)
      if ((st = forward_on(p, replay(m), replay(m + 1), vel, n, scratch, s)))
        return st;
    for (int m = len - 1; m >= 0; --m) {
      float *x = replay(m);
      const float *P[3] = {x, x + n, x + 2 * n};
      float *li = lam[cur], *lo = lam[cur ^ 1];
      const float *LXN[3] = {li, li + n, li + 2 * n};
      const float *LVN[3] = {li + n3, li + n3 + n, li + n3 + 2 * n};
      float *LX[3] = {lo, lo + n, lo + 2 * n};
      float *LV[3] = {lo + n3, lo + n3 + n, lo + n3 + 2 * n};
      float *none[3] = {nullptr, nullptr, nullptr};
      if ((st = launch_adjoint(c, skip, false, cfg_a, P, LXN, LVN, LX, LV,
                               none, none, (int)n, scratch, s)))
        return st;
      cur ^= 1;
    }
  }
  return copy_d2d(lam0, lam[cur], 2 * n3, s);
}

} // extern "C"
