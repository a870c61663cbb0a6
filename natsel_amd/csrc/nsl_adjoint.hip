// Adjoint of a recorded batched rollout (the differentiable model of nsl_diff.h), and the
// same model in double precision for finite differences.
//
// k_adjoint: one wavefront per world walks the recorded steps backwards. Each step is first
// recomputed from its recorded start state with the model's float arithmetic (bit-identical
// to the step kernels; the recompute is compared with the next recorded state), saving the
// body states and impulses at the start of every colour phase of every sub-step; then the
// stages are reversed phase by phase from those checkpoints. Across steps the adjoint
// carries the body velocities, the centre offsets (through the separations and the final
// centres) and the impulses (through the warm-start links).
//
// k_model_loss: one wavefront per case runs the model forward in double precision from the
// recorded first state with one parameter perturbed, and returns the loss.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "nsl_batch.h"
#include "nsl_diff.h"

#define ADJ_CHECK(expr)                                                                          \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace nsl {
namespace {

namespace df = diff;

struct ModelRows {
  df::PushRows contact, fixed;  // dynamic-dynamic and static contact softness
  float h, inv_h;
  int32_t sub_steps;
  Vec3 gravity;
};

ModelRows model_rows(const StepParams& p) {
  ModelRows m;
  m.contact = {p.contact_softness.mass_scale * p.contact_softness.bias_rate, p.contact_softness.mass_scale,
               p.contact_softness.impulse_scale, p.inv_h, -p.contact_speed};
  m.fixed = {p.static_softness.mass_scale * p.static_softness.bias_rate, p.static_softness.mass_scale,
             p.static_softness.impulse_scale, p.inv_h, -p.contact_speed};
  m.h = p.h;
  m.inv_h = p.inv_h;
  m.sub_steps = p.sub_steps;
  m.gravity = p.gravity;
  return m;
}

struct AdjointArgs {
  BatchDevice b;
  BatchTape t;
  ModelRows rows;
  const float* targets;  // [worlds][3]
  float* loss;           // [worlds]
  float* grad_v;         // [awake][3]
  float* grad_w;         // [awake][3]
  float* grad_f;         // [awake]
  int32_t* status;       // [worlds]
  df::Body<float>* ck_bodies;     // [worlds][n_ckpt][kBatchLanes]
  df::Impulses<float>* ck_imp;    // [worlds][n_ckpt][kBatchLanes]
  df::Impulses<float>* carry;     // [worlds][kBatchLanes]: adjoint of the final impulses of the step being reversed
  int32_t n_ckpt;                 // per step: sub_steps * (3 * max_colours + 1)
  int32_t max_colours;
};

__device__ __forceinline__ df::Impulses<float> warm_values(const Constraint& c) {
  df::Impulses<float> im;
  for (int i = 0; i < 4; ++i) im.ni[i] = i < c.point_count ? c.points[i].normal_impulse : 0.0f;
  im.fx = c.friction_impulse.x;
  im.fy = c.friction_impulse.y;
  im.ti = c.twist_impulse;
  return im;
}

__device__ __forceinline__ int colour_of(const int32_t* col, int slot) {
  for (int a = 0; a < col[0]; ++a)
    if (slot >= col[1 + a] && slot < col[1 + kOverflowColor + a]) return a;
  return -1;
}

__global__ __launch_bounds__(kBatchLanes) void k_adjoint(AdjointArgs A) {
  __shared__ df::Body<float> S[kBatchLanes];
  __shared__ df::BodyAdj<float> SB[kBatchLanes];
  __shared__ df::Impulses<float> I[kBatchLanes];
  __shared__ df::Impulses<float> IB[kBatchLanes];
  __shared__ float CBX[kBatchLanes], CBY[kBatchLanes], CBZ[kBatchLanes];
  __shared__ int32_t col[kTapeColourWords];
  __shared__ int s_status;
  const int world = blockIdx.x;
  const int lane = threadIdx.x;
  const BatchDevice& b = A.b;
  const BatchTape& t = A.t;
  const BatchWorldDesc wd = b.worlds[world];
  const int nb = wd.n_awake;
  const int W = t.n_worlds;
  if (lane == 0) {
    int st = (nb > t.max_bodies || nb > kBatchLanes) ? 1 : 0;
    for (int k = 0; k < t.steps && st == 0; ++k)
      if (t.count[(size_t)k * W + world] < 0 || t.count[(size_t)k * W + world] > kBatchLanes) st = 1;
    s_status = st;
  }
  __syncthreads();
  if (s_status != 0) {
    if (lane == 0) {
      A.loss[world] = __builtin_nanf("");
      A.status[world] = s_status;
    }
    return;
  }
  const ModelRows& R = A.rows;
  // loss and its adjoint at the final centres
  const BodySim* sims = b.sims + wd.awake_off;
  double m = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
  const size_t wT = (size_t)t.steps * W + world;
  if (lane < nb) {
    const float im = sims[lane].inv_mass;
    m = im > 0.0f ? 1.0 / (double)im : 0.0;
    const Pos c = t.centers[wT * t.max_bodies + lane];
    mx = m * c.x;
    my = m * c.y;
    mz = m * c.z;
  }
  double M = m, SX = mx, SY = my, SZ = mz;
  for (int off = 32; off > 0; off >>= 1) {
    M += __shfl_xor(M, off, 64);
    SX += __shfl_xor(SX, off, 64);
    SY += __shfl_xor(SY, off, 64);
    SZ += __shfl_xor(SZ, off, 64);
  }
  const double ex = SX / M - (double)A.targets[3 * world + 0];
  const double ey = SY / M - (double)A.targets[3 * world + 1];
  const double ez = SZ / M - (double)A.targets[3 * world + 2];
  if (lane == 0) A.loss[world] = (float)(ex * ex + ey * ey + ez * ez);
  if (lane < kBatchLanes) {
    const float k = (float)(2.0 * m / M);
    CBX[lane] = k * (float)ex;
    CBY[lane] = k * (float)ey;
    CBZ[lane] = k * (float)ez;
    SB[lane].v = SB[lane].w = SB[lane].dp = df::zero<float>();
    SB[lane].dq = {df::zero<float>(), 0.0f};
    df::Impulses<float> z = {{0, 0, 0, 0}, 0, 0, 0};
    A.carry[(size_t)world * kBatchLanes + lane] = z;
  }
  if (lane == 0) s_status = 0;
  __syncthreads();

  df::Body<float>* ckb = A.ck_bodies + (size_t)world * A.n_ckpt * kBatchLanes;
  df::Impulses<float>* cki = A.ck_imp + (size_t)world * A.n_ckpt * kBatchLanes;
  const int na_max = A.max_colours;
  const int per_sub = 3 * na_max + 1;

  for (int k = t.steps - 1; k >= 0; --k) {
    const size_t ws = (size_t)k * W + world;
    const int n = t.count[ws];
    const Constraint* C = t.cons + ws * t.max_con;
    if (lane < kTapeColourWords) col[lane] = t.colours[ws * kTapeColourWords + lane];
    if (lane < nb) {
      const BodyState s = t.states[ws * t.max_bodies + lane];
      S[lane] = {df::mk<float>(s.v), df::mk<float>(s.w), df::zero<float>(), {df::zero<float>(), 1.0f}, s.flags};
    }
    __syncthreads();
    const int na = col[0];
    const int my_colour = lane < n ? colour_of(col, lane) : -1;
    df::ConParams<float> cp;
    df::Impulses<float> warm = {{0, 0, 0, 0}, 0, 0, 0};
    if (lane < n) {
      const Constraint& c = C[lane];
      for (int i = 0; i < 4; ++i) cp.bs[i] = i < c.point_count ? c.points[i].base_separation : 0.0f;
      cp.mu = c.friction;
      warm = warm_values(c);
      I[lane] = warm;
    }
    __syncthreads();
    // ---- forward recompute with checkpoints
    for (int sub = 0; sub < R.sub_steps; ++sub) {
      const int base = sub * per_sub;
      uint8_t* br = t.branches + ((ws * t.sub_steps + sub) * t.max_con + (lane < t.max_con ? lane : 0)) * kTapeBranchBytes;
      if (lane < nb) {
        const BodySim& bs = sims[lane];
        df::integrate_velocities(S[lane], R.h, bs.inv_mass, bs.gravity_scale, bs.linear_damping, bs.angular_damping, R.gravity);
      }
      __syncthreads();
      for (int a = 0; a < na; ++a) {
        ckb[(size_t)(base + a) * kBatchLanes + lane] = S[lane];
        cki[(size_t)(base + a) * kBatchLanes + lane] = I[lane];
        __syncthreads();
        if (my_colour == a) df::warm_start(C[lane], I[lane], S);
        __syncthreads();
      }
      for (int a = 0; a < na; ++a) {
        ckb[(size_t)(base + na_max + a) * kBatchLanes + lane] = S[lane];
        cki[(size_t)(base + na_max + a) * kBatchLanes + lane] = I[lane];
        __syncthreads();
        if (my_colour == a) df::push(C[lane], cp, I[lane], S, C[lane].is_static ? R.fixed : R.contact, br);
        __syncthreads();
      }
      ckb[(size_t)(base + 2 * na_max) * kBatchLanes + lane] = S[lane];
      __syncthreads();
      if (lane < nb) df::integrate_positions(S[lane], R.h);
      __syncthreads();
      for (int a = 0; a < na; ++a) {
        ckb[(size_t)(base + 2 * na_max + 1 + a) * kBatchLanes + lane] = S[lane];
        cki[(size_t)(base + 2 * na_max + 1 + a) * kBatchLanes + lane] = I[lane];
        __syncthreads();
        if (my_colour == a) df::relax(C[lane], cp, I[lane], S, R.inv_h, br + 4);
        __syncthreads();
      }
    }
    // the recompute must land on the recorded next state
    if (lane < nb) {
      const BodyState s = t.states[(ws + W) * t.max_bodies + lane];
      const df::Body<float>& x = S[lane];
      if (x.v.x != s.v.x || x.v.y != s.v.y || x.v.z != s.v.z || x.w.x != s.w.x || x.w.y != s.w.y || x.w.z != s.w.z)
        atomicOr(&s_status, 2);
    }
    // ---- reverse
    if (lane < nb) {
      SB[lane].dp = {CBX[lane], CBY[lane], CBZ[lane]};  // centre_{k+1} = centre_k + dp
      SB[lane].dq = {df::zero<float>(), 0.0f};
    }
    df::ConParamsAdj<float> cb = {{0, 0, 0, 0}, 0};
    if (lane < kBatchLanes) IB[lane] = A.carry[(size_t)world * kBatchLanes + lane];
    __syncthreads();
    for (int sub = R.sub_steps - 1; sub >= 0; --sub) {
      const int base = sub * per_sub;
      for (int a = na - 1; a >= 0; --a) {
        S[lane] = ckb[(size_t)(base + 2 * na_max + 1 + a) * kBatchLanes + lane];
        I[lane] = cki[(size_t)(base + 2 * na_max + 1 + a) * kBatchLanes + lane];
        __syncthreads();
        if (my_colour == a) df::relax_adj(C[lane], cp, I[lane], S, SB, IB[lane], cb, R.inv_h);
        __syncthreads();
      }
      S[lane] = ckb[(size_t)(base + 2 * na_max) * kBatchLanes + lane];
      __syncthreads();
      if (lane < nb && (S[lane].flags & bflag::kDynamic)) df::integrate_positions_adj(S[lane], SB[lane], R.h);
      __syncthreads();
      for (int a = na - 1; a >= 0; --a) {
        S[lane] = ckb[(size_t)(base + na_max + a) * kBatchLanes + lane];
        I[lane] = cki[(size_t)(base + na_max + a) * kBatchLanes + lane];
        __syncthreads();
        if (my_colour == a) df::push_adj(C[lane], cp, I[lane], S, SB, IB[lane], cb, C[lane].is_static ? R.fixed : R.contact);
        __syncthreads();
      }
      for (int a = na - 1; a >= 0; --a) {
        S[lane] = ckb[(size_t)(base + a) * kBatchLanes + lane];
        __syncthreads();
        if (my_colour == a) df::warm_start_adj(C[lane], S, SB, IB[lane]);
        __syncthreads();
      }
      if (lane < nb) {
        const BodySim& bs = sims[lane];
        df::integrate_velocities_adj(SB[lane], R.h, bs.linear_damping, bs.angular_damping);
      }
      __syncthreads();
    }
    // the step starts from zero displacement and identity rotation: their adjoints end here
    if (lane < nb) {
      SB[lane].dp = df::zero<float>();
      SB[lane].dq = {df::zero<float>(), 0.0f};
    }
    // separations follow the centres: bs = recorded + n . (dc_b - dc_a)
    if (lane < n) {
      const Constraint& c = C[lane];
      float sb = 0.0f;
      for (int i = 0; i < c.point_count; ++i) sb += cb.bs[i];
      if (sb != 0.0f) {
        const Vec3 nn = c.normal;
        if (c.index_b >= 0 && (S[c.index_b].flags & bflag::kDynamic)) {
          atomicAdd(&CBX[c.index_b], sb * nn.x);
          atomicAdd(&CBY[c.index_b], sb * nn.y);
          atomicAdd(&CBZ[c.index_b], sb * nn.z);
        }
        if (c.index_a >= 0 && (S[c.index_a].flags & bflag::kDynamic)) {
          atomicAdd(&CBX[c.index_a], -sb * nn.x);
          atomicAdd(&CBY[c.index_a], -sb * nn.y);
          atomicAdd(&CBZ[c.index_a], -sb * nn.z);
        }
      }
      // friction = sqrt(f_a f_b) of the two shapes
      if (cb.mu != 0.0f && c.friction > 0.0f) {
        const uint32_t pair = t.pairs[ws * t.max_con + lane];
        const Shape& sa = b.shapes[wd.shape_off + (int)(pair >> 16)];
        const Shape& sbs = b.shapes[wd.shape_off + (int)(pair & 0xFFFFu)];
        if (sa.body_enc >= 0) atomicAdd(&A.grad_f[wd.awake_off + sa.body_enc], cb.mu * sbs.friction / (2.0f * c.friction));
        if (sbs.body_enc >= 0) atomicAdd(&A.grad_f[wd.awake_off + sbs.body_enc], cb.mu * sa.friction / (2.0f * c.friction));
      }
    }
    __syncthreads();
    // warm-start impulses came from the previous step's final impulses
    df::Impulses<float>* carry = A.carry + (size_t)world * kBatchLanes;
    if (lane < kBatchLanes) {
      df::Impulses<float> z = {{0, 0, 0, 0}, 0, 0, 0};
      carry[lane] = z;
    }
    __syncthreads();
    if (k > 0 && lane < n) {
      const int32_t link = t.links[ws * t.max_con + lane];
      if (link >= 0) {
        const Constraint& c = C[lane];
        const Constraint& cprev = t.cons[(ws - W) * t.max_con + link];
        df::Impulses<float>& dst = carry[link];
        for (int i = 0; i < c.point_count; ++i) {
          const uint8_t q = t.pmap[(ws * t.max_con + lane) * 4 + i];
          if (q != kTapeNoPoint) dst.ni[q] += IB[lane].ni[i];
        }
        dst.ti += IB[lane].ti;
        // fx = F . t1, fy = F . t2 with F = fx' t1' + fy' t2'
        const df::V3<float> fb = df::add(df::scale(IB[lane].fx, df::mk<float>(c.tangent1)), df::scale(IB[lane].fy, df::mk<float>(c.tangent2)));
        dst.fx += df::dot(fb, df::mk<float>(cprev.tangent1));
        dst.fy += df::dot(fb, df::mk<float>(cprev.tangent2));
      }
    }
    __syncthreads();
  }
  if (lane < nb) {
    const size_t g = (size_t)wd.awake_off + lane;
    A.grad_v[3 * g + 0] = SB[lane].v.x;
    A.grad_v[3 * g + 1] = SB[lane].v.y;
    A.grad_v[3 * g + 2] = SB[lane].v.z;
    A.grad_w[3 * g + 0] = SB[lane].w.x;
    A.grad_w[3 * g + 1] = SB[lane].w.y;
    A.grad_w[3 * g + 2] = SB[lane].w.z;
  }
  if (lane == 0) A.status[world] = s_status;
}

// ---------------------------------------------------------------------------------------
// the model in double precision, one case (world, perturbed parameter) per wavefront

struct ModelArgs {
  BatchDevice b;
  BatchTape t;
  ModelRows rows;
  const int32_t* worlds;
  const int32_t* params;
  const double* deltas;
  const float* targets;
  double* out;
  int32_t frozen;  // take every branch from the adjoint's record
};

__global__ __launch_bounds__(kBatchLanes) void k_model_loss(ModelArgs A) {
  __shared__ df::Body<double> S[kBatchLanes];
  __shared__ df::Impulses<double> F[kBatchLanes];  // final impulses of the previous step
  __shared__ double DCX[kBatchLanes], DCY[kBatchLanes], DCZ[kBatchLanes], FR[kBatchLanes];
  __shared__ int32_t col[kTapeColourWords];
  const int cs = blockIdx.x;
  const int lane = threadIdx.x;
  const int world = A.worlds[cs];
  const int param = A.params[cs];
  const double delta = A.deltas[cs];
  const BatchDevice& b = A.b;
  const BatchTape& t = A.t;
  const BatchWorldDesc wd = b.worlds[world];
  const int nb = wd.n_awake;
  const int W = t.n_worlds;
  const ModelRows& R = A.rows;
  const BodySim* sims = b.sims + wd.awake_off;
  const int pb = param / 7, pk = param % 7;
  if (lane < nb) {
    const BodyState s = t.states[(size_t)world * t.max_bodies + lane];
    df::Body<double> x = {df::mk<double>(s.v), df::mk<double>(s.w), df::zero<double>(), {df::zero<double>(), 1.0}, s.flags};
    if (lane == pb && pk < 3) (&x.v.x)[pk] += delta;
    if (lane == pb && pk >= 3 && pk < 6) (&x.w.x)[pk - 3] += delta;
    S[lane] = x;
    DCX[lane] = DCY[lane] = DCZ[lane] = 0.0;
  }
  // friction parameter of each awake body's (first) shape
  if (lane < nb) FR[lane] = 0.0;
  __syncthreads();
  if (lane == 0)
    for (int s = 0; s < wd.n_shapes; ++s) {
      const Shape& sh = b.shapes[wd.shape_off + s];
      if (sh.body_enc >= 0) FR[sh.body_enc] = (double)sh.friction + ((sh.body_enc == pb && pk == 6) ? delta : 0.0);
    }
  __syncthreads();
  for (int k = 0; k < t.steps; ++k) {
    const size_t ws = (size_t)k * W + world;
    const int n = t.count[ws];
    const Constraint* C = t.cons + ws * t.max_con;
    if (lane < kTapeColourWords) col[lane] = t.colours[ws * kTapeColourWords + lane];
    df::ConParams<double> cp;
    df::Impulses<double> im = {{0, 0, 0, 0}, 0, 0, 0};
    if (lane < n) {
      const Constraint& c = C[lane];
      const int32_t link = k > 0 ? t.links[ws * t.max_con + lane] : -1;
      for (int i = 0; i < 4; ++i) {
        double ni = i < c.point_count ? (double)c.points[i].normal_impulse : 0.0;
        if (link >= 0 && i < c.point_count) {
          const uint8_t q = t.pmap[(ws * t.max_con + lane) * 4 + i];
          if (q != kTapeNoPoint) ni = F[link].ni[q];
        }
        im.ni[i] = ni;
      }
      if (link >= 0) {
        const Constraint& cprev = t.cons[(ws - W) * t.max_con + link];
        const df::V3<double> fv = df::add(df::scale(F[link].fx, df::mk<double>(cprev.tangent1)),
                                          df::scale(F[link].fy, df::mk<double>(cprev.tangent2)));
        im.fx = df::dot(fv, df::mk<double>(c.tangent1));
        im.fy = df::dot(fv, df::mk<double>(c.tangent2));
        im.ti = F[link].ti;
      } else {
        im.fx = c.friction_impulse.x;
        im.fy = c.friction_impulse.y;
        im.ti = c.twist_impulse;
      }
      const df::V3<double> nn = df::mk<double>(c.normal);
      df::V3<double> dca = df::zero<double>(), dcb = df::zero<double>();
      if (c.index_a >= 0) dca = {DCX[c.index_a], DCY[c.index_a], DCZ[c.index_a]};
      if (c.index_b >= 0) dcb = {DCX[c.index_b], DCY[c.index_b], DCZ[c.index_b]};
      const double shift = df::dot(nn, df::sub(dcb, dca));
      for (int i = 0; i < 4; ++i) cp.bs[i] = (i < c.point_count ? (double)c.points[i].base_separation : 0.0) + shift;
      const uint32_t pair = t.pairs[ws * t.max_con + lane];
      const Shape& sa = b.shapes[wd.shape_off + (int)(pair >> 16)];
      const Shape& sbs = b.shapes[wd.shape_off + (int)(pair & 0xFFFFu)];
      const double fa = sa.body_enc >= 0 ? FR[sa.body_enc] : (double)sa.friction;
      const double fb = sbs.body_enc >= 0 ? FR[sbs.body_enc] : (double)sbs.friction;
      cp.mu = sqrt(fa * fb);
    }
    __shared__ df::Impulses<double> I[kBatchLanes];
    if (lane < n) I[lane] = im;
    if (lane < nb) {
      S[lane].dp = df::zero<double>();
      S[lane].dq = {df::zero<double>(), 1.0};
    }
    __syncthreads();
    const int na = col[0];
    const int my_colour = lane < n ? colour_of(col, lane) : -1;
    for (int sub = 0; sub < R.sub_steps; ++sub) {
      const uint8_t* br =
          A.frozen ? t.branches + ((ws * t.sub_steps + sub) * t.max_con + (lane < t.max_con ? lane : 0)) * kTapeBranchBytes : nullptr;
      if (lane < nb) {
        const BodySim& bs = sims[lane];
        df::integrate_velocities(S[lane], R.h, bs.inv_mass, bs.gravity_scale, bs.linear_damping, bs.angular_damping, R.gravity);
      }
      __syncthreads();
      for (int a = 0; a < na; ++a) {
        if (my_colour == a) df::warm_start(C[lane], I[lane], S);
        __syncthreads();
      }
      for (int a = 0; a < na; ++a) {
        if (my_colour == a) df::push(C[lane], cp, I[lane], S, C[lane].is_static ? R.fixed : R.contact, (uint8_t*)nullptr, br);
        __syncthreads();
      }
      if (lane < nb) df::integrate_positions(S[lane], R.h);
      __syncthreads();
      for (int a = 0; a < na; ++a) {
        if (my_colour == a) df::relax(C[lane], cp, I[lane], S, R.inv_h, (uint8_t*)nullptr, br ? br + 4 : nullptr);
        __syncthreads();
      }
    }
    if (lane < n) F[lane] = I[lane];
    if (lane < nb) {
      // centre_{k+1} = centre_k + dp; the offset from the recording moves by dp - recorded dp
      const Pos c0 = t.centers[ws * t.max_bodies + lane], c1 = t.centers[(ws + W) * t.max_bodies + lane];
      DCX[lane] += S[lane].dp.x - (c1.x - c0.x);
      DCY[lane] += S[lane].dp.y - (c1.y - c0.y);
      DCZ[lane] += S[lane].dp.z - (c1.z - c0.z);
    }
    __syncthreads();
  }
  double m = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
  if (lane < nb) {
    const float imv = sims[lane].inv_mass;
    m = imv > 0.0f ? 1.0 / (double)imv : 0.0;
    const Pos c = t.centers[((size_t)t.steps * W + world) * t.max_bodies + lane];
    mx = m * (c.x + DCX[lane]);
    my = m * (c.y + DCY[lane]);
    mz = m * (c.z + DCZ[lane]);
  }
  for (int off = 32; off > 0; off >>= 1) {
    m += __shfl_xor(m, off, 64);
    mx += __shfl_xor(mx, off, 64);
    my += __shfl_xor(my, off, 64);
    mz += __shfl_xor(mz, off, 64);
  }
  if (lane == 0) {
    const double ex = mx / m - (double)A.targets[3 * world + 0];
    const double ey = my / m - (double)A.targets[3 * world + 1];
    const double ez = mz / m - (double)A.targets[3 * world + 2];
    A.out[cs] = ex * ex + ey * ey + ez * ez;
  }
}

template <class T> T* tape_alloc(std::vector<void*>& list, size_t n) {
  T* p = nullptr;
  ADJ_CHECK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  list.push_back(p);
  return p;
}

}  // namespace

void BatchStepper::free_tape() {
  for (void* p : tape_allocs_) (void)hipFree(p);
  tape_allocs_.clear();
  tape_ = BatchTape{};
}

void BatchStepper::record(float dt, int32_t sub_steps, int32_t steps, int32_t max_constraints) {
  if (!(dt > 0.0f) || steps <= 0) throw std::invalid_argument("recording needs a positive time step and step count");
  if (max_constraints <= 0 || max_constraints > kBatchLanes)
    throw std::invalid_argument("the tape holds 1.." + std::to_string(kBatchLanes) + " constraints per world");
  ADJ_CHECK(hipSetDevice(device_));
  if (!uploaded_) upload();
  ADJ_CHECK(hipStreamSynchronize(stream_));
  free_tape();
  const int W = d_.n_worlds;
  BatchTape t{};
  t.max_con = max_constraints;
  t.max_bodies = kBatchLanes;
  t.steps = steps;
  t.n_worlds = W;
  const size_t cells = (size_t)steps * W;
  t.cons = tape_alloc<Constraint>(tape_allocs_, cells * t.max_con);
  t.features = tape_alloc<uint32_t>(tape_allocs_, cells * t.max_con * 4);
  t.pairs = tape_alloc<uint32_t>(tape_allocs_, cells * t.max_con);
  t.links = tape_alloc<int32_t>(tape_allocs_, cells * t.max_con);
  t.pmap = tape_alloc<uint8_t>(tape_allocs_, cells * t.max_con * 4);
  t.count = tape_alloc<int32_t>(tape_allocs_, cells);
  t.colours = tape_alloc<int32_t>(tape_allocs_, cells * kTapeColourWords);
  t.states = tape_alloc<BodyState>(tape_allocs_, (cells + W) * t.max_bodies);
  t.centers = tape_alloc<Pos>(tape_allocs_, (cells + W) * t.max_bodies);
  t.sub_steps = std::max(1, sub_steps);
  t.branches = tape_alloc<uint8_t>(tape_allocs_, cells * t.sub_steps * t.max_con * kTapeBranchBytes);
  tape_branches_ = false;
  ADJ_CHECK(hipMemsetAsync(t.count, 0xFF, cells * sizeof(int32_t), stream_));  // -1: not recorded
  ADJ_CHECK(hipMemsetAsync(t.colours, 0, cells * kTapeColourWords * sizeof(int32_t), stream_));
  tape_ = t;
  tape_params_ = step_params(dt, sub_steps);
  BatchDevice b = d_;
  b.tape = t;
  b.tape_step = 0;
  ADJ_CHECK(hipEventRecord(ev0_, stream_));
  launch_batch_step(b, tape_params_, steps, stream_);
  ADJ_CHECK(hipEventRecord(ev1_, stream_));
  ADJ_CHECK(hipEventSynchronize(ev1_));
  ADJ_CHECK(hipEventElapsedTime(&record_ms_, ev0_, ev1_));
  body_steps_ += total_awake_ * (int64_t)steps;
  device_newer_ = true;
  // the largest colour count sizes the adjoint's checkpoints
  std::vector<int32_t> cols(cells * kTapeColourWords);
  ADJ_CHECK(hipMemcpy(cols.data(), t.colours, cols.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
  tape_max_colours_ = 1;
  for (size_t i = 0; i < cells; ++i) tape_max_colours_ = std::max(tape_max_colours_, cols[i * kTapeColourWords]);
}

void BatchStepper::adjoint(const float* targets, float* loss, float* grad_v, float* grad_w, float* grad_friction, int32_t* status) {
  if (!tape_.cons) throw std::invalid_argument("no recorded rollout");
  ADJ_CHECK(hipSetDevice(device_));
  const int W = d_.n_worlds;
  std::vector<void*> tmp;
  AdjointArgs a{};
  a.b = d_;
  a.t = tape_;
  a.rows = model_rows(tape_params_);
  float* d_targets = tape_alloc<float>(tmp, (size_t)W * 3);
  a.targets = d_targets;
  a.loss = tape_alloc<float>(tmp, W);
  a.grad_v = tape_alloc<float>(tmp, (size_t)total_awake_ * 3);
  a.grad_w = tape_alloc<float>(tmp, (size_t)total_awake_ * 3);
  a.grad_f = tape_alloc<float>(tmp, (size_t)total_awake_);
  a.status = tape_alloc<int32_t>(tmp, W);
  a.max_colours = tape_max_colours_;
  a.n_ckpt = tape_params_.sub_steps * (3 * a.max_colours + 1);
  a.ck_bodies = tape_alloc<diff::Body<float>>(tmp, (size_t)W * a.n_ckpt * kBatchLanes);
  a.ck_imp = tape_alloc<diff::Impulses<float>>(tmp, (size_t)W * a.n_ckpt * kBatchLanes);
  a.carry = tape_alloc<diff::Impulses<float>>(tmp, (size_t)W * kBatchLanes);
  try {
    ADJ_CHECK(hipMemcpyAsync(d_targets, targets, (size_t)W * 3 * sizeof(float), hipMemcpyHostToDevice, stream_));
    ADJ_CHECK(hipMemsetAsync(a.grad_f, 0, (size_t)total_awake_ * sizeof(float), stream_));
    ADJ_CHECK(hipMemsetAsync(a.ck_bodies, 0, (size_t)W * a.n_ckpt * kBatchLanes * sizeof(diff::Body<float>), stream_));
    ADJ_CHECK(hipMemsetAsync(a.ck_imp, 0, (size_t)W * a.n_ckpt * kBatchLanes * sizeof(diff::Impulses<float>), stream_));
    ADJ_CHECK(hipEventRecord(ev0_, stream_));
    hipLaunchKernelGGL(k_adjoint, dim3(W), dim3(kBatchLanes), 0, stream_, a);
    ADJ_CHECK(hipGetLastError());
    ADJ_CHECK(hipEventRecord(ev1_, stream_));
    ADJ_CHECK(hipEventSynchronize(ev1_));
    ADJ_CHECK(hipEventElapsedTime(&adjoint_ms_, ev0_, ev1_));
    tape_branches_ = true;
    if (loss) ADJ_CHECK(hipMemcpy(loss, a.loss, W * sizeof(float), hipMemcpyDeviceToHost));
    if (grad_v) ADJ_CHECK(hipMemcpy(grad_v, a.grad_v, (size_t)total_awake_ * 3 * sizeof(float), hipMemcpyDeviceToHost));
    if (grad_w) ADJ_CHECK(hipMemcpy(grad_w, a.grad_w, (size_t)total_awake_ * 3 * sizeof(float), hipMemcpyDeviceToHost));
    if (grad_friction) ADJ_CHECK(hipMemcpy(grad_friction, a.grad_f, (size_t)total_awake_ * sizeof(float), hipMemcpyDeviceToHost));
    if (status) ADJ_CHECK(hipMemcpy(status, a.status, W * sizeof(int32_t), hipMemcpyDeviceToHost));
  } catch (...) {
    for (void* p : tmp) (void)hipFree(p);
    throw;
  }
  for (void* p : tmp) (void)hipFree(p);
}

void BatchStepper::model_losses(const int32_t* worlds, const int32_t* params, const double* deltas, int32_t n, const float* targets,
                                double* out, bool frozen) {
  if (!tape_.cons) throw std::invalid_argument("no recorded rollout");
  if (frozen && !tape_branches_) throw std::invalid_argument("the frozen model needs the branch record of adjoint()");
  if (n <= 0) return;
  ADJ_CHECK(hipSetDevice(device_));
  const int W = d_.n_worlds;
  for (int32_t i = 0; i < n; ++i) {
    if (worlds[i] < 0 || worlds[i] >= W) throw std::invalid_argument("model_losses: no such world");
    if (params[i] < 0 || params[i] / 7 >= desc_[worlds[i]].n_awake) throw std::invalid_argument("model_losses: no such parameter");
  }
  std::vector<void*> tmp;
  ModelArgs m{};
  m.b = d_;
  m.t = tape_;
  m.rows = model_rows(tape_params_);
  int32_t* d_worlds = tape_alloc<int32_t>(tmp, n);
  int32_t* d_params = tape_alloc<int32_t>(tmp, n);
  double* d_deltas = tape_alloc<double>(tmp, n);
  float* d_targets = tape_alloc<float>(tmp, (size_t)W * 3);
  m.worlds = d_worlds;
  m.params = d_params;
  m.deltas = d_deltas;
  m.targets = d_targets;
  m.out = tape_alloc<double>(tmp, n);
  m.frozen = frozen ? 1 : 0;
  try {
    ADJ_CHECK(hipMemcpyAsync(d_worlds, worlds, n * sizeof(int32_t), hipMemcpyHostToDevice, stream_));
    ADJ_CHECK(hipMemcpyAsync(d_params, params, n * sizeof(int32_t), hipMemcpyHostToDevice, stream_));
    ADJ_CHECK(hipMemcpyAsync(d_deltas, deltas, n * sizeof(double), hipMemcpyHostToDevice, stream_));
    ADJ_CHECK(hipMemcpyAsync(d_targets, targets, (size_t)W * 3 * sizeof(float), hipMemcpyHostToDevice, stream_));
    hipLaunchKernelGGL(k_model_loss, dim3(n), dim3(kBatchLanes), 0, stream_, m);
    ADJ_CHECK(hipGetLastError());
    ADJ_CHECK(hipMemcpyAsync(out, m.out, n * sizeof(double), hipMemcpyDeviceToHost, stream_));
    ADJ_CHECK(hipStreamSynchronize(stream_));
  } catch (...) {
    for (void* p : tmp) (void)hipFree(p);
    throw;
  }
  for (void* p : tmp) (void)hipFree(p);
}

}  // namespace nsl
