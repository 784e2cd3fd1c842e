// pbd.hip -- ClaySolver: shape-matched particle bodies against a TracedField
// (lib/src/sim/pbd.ts:78-424) over a ParticleSet (lib/src/sim/particles.ts:18-206).
//
// Per substep four launches, as the reference dispatches them (pbd.ts:400-415):
//   k_pbd_predict   per particle: BDF2 predictor, field collision, positional friction
//   k_pbd_match     per body (256 lanes): centroids, covariance, polar rotation
//   k_pbd_apply     per particle: shape matching, collision projection, plasticity
//   k_pbd_finalize  per particle: velocity from positions, speed cap
// plus k_pbd_kick once before the substeps when an impulse is pending.
//
// Particles are structure-of-arrays planes (pos, prev, vel, velPrev, rest) and an
// owner per particle; bodies are one record each.  Every expression follows the
// reference's operation order with no contraction (this file is built with
// -ffp-contract=off), and the match reductions keep the reference's lane-strided
// sums and pairwise tree, so results are deterministic and reproducible on the
// CPU.  sin and cos (quatFromAxisAngle) come from one polynomial of this file,
// defined below, so that they too are reproducible.
#include "clay_internal.h"

namespace clay {

namespace {

constexpr int kReduce = 256;  // lanes per body in the match pass (pbd.ts:8)
constexpr int kGroup = 256;   // threads per workgroup of the per-particle passes

struct PbdK {
  float g[3];
  float h;  // substep length
  float radius, friction, damping, max_speed;
  float push[3];
  uint32_t push_body;
  float kick[3];
  uint32_t capacity, max_bodies;
  int32_t rot_iters;
};

struct f4 {
  float x, y, z, w;
};

// Mirrors clay_body_desc, plus the state the match pass maintains.
struct BodyDev {
  uint32_t first, count, flags;
  float stiffness, plasticity;
  float com[3], rest_com[3], rot[4];
};

struct Parts {
  float* p[3];  // pos
  float* q[3];  // prev
  float* v[3];  // vel
  float* u[3];  // velPrev
  float* r[3];  // rest
  const uint32_t* body;
};

__device__ __forceinline__ f3 ld(float* const a[3], uint32_t i) { return mk(a[0][i], a[1][i], a[2][i]); }
__device__ __forceinline__ void st(float* const a[3], uint32_t i, f3 v) {
  a[0][i] = v.x;
  a[1][i] = v.y;
  a[2][i] = v.z;
}

// sin and cos of x: x = k pi/2 + r with k = floor(x 2/pi + 1/2) and r reduced in two
// steps (1.5703125 has 8 significant bits, so k * 1.5703125 is exact for |k| < 2^16),
// then the Taylor series of sin to r^9 and of cos to r^10 on |r| <= pi/4 (truncation
// below 1.2e-10), Horner form, and the quadrant's swap and signs.
__device__ inline void sin_cos(float x, float* s, float* c) {
  const float k = floorf(x * 0.636619772f + 0.5f);
  const float r = (x - k * 1.5703125f) - k * 4.83826794897e-4f;
  const float r2 = r * r;
  const float sp = r + r * r2 * (-0.166666667f + r2 * (8.33333333e-3f + r2 * (-1.98412698e-4f + r2 * 2.75573192e-6f)));
  const float cp =
      1.0f + r2 * (-0.5f + r2 * (4.16666667e-2f + r2 * (-1.38888889e-3f + r2 * (2.48015873e-5f - r2 * 2.75573192e-7f))));
  const int q = ((int)k) & 3;
  const float ss = (q & 1) ? cp : sp, cc = (q & 1) ? sp : cp;
  *s = (q == 2 || q == 3) ? -ss : ss;
  *c = (q == 1 || q == 2) ? -cc : cc;
}

// math/gpu.ts:5-14
__device__ __forceinline__ f4 quat_mul(f4 a, f4 b) {
  return f4{a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
            a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
// math/gpu.ts:15-25: v + t w + cross(q.xyz, t) with t = 2 cross(q.xyz, v)
__device__ __forceinline__ f3 quat_rotate(f4 q, f3 v) {
  const f3 qv = mk(q.x, q.y, q.z);
  const f3 t = cross3(qv, v) * 2.0f;
  return v + t * q.w + cross3(qv, t);
}
__device__ __forceinline__ f4 quat_conj(f4 q) { return f4{-q.x, -q.y, -q.z, q.w}; }

// extractRotation (math/gpu.ts:55-78; Mueller et al. 2016): warm-started
// quaternion iterations towards the rotation of the polar decomposition of A
// (columns a0, a1, a2).  normalize is v / length(v).
__device__ f4 extract_rotation(f3 a0, f3 a1, f3 a2, f4 q, int iterations) {
  for (int i = 0; i < iterations; i++) {
    const float x = q.x, y = q.y, z = q.z, w = q.w;
    // quatToMat3 (math/gpu.ts:32-45), columns
    const f3 c0 = mk(1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y + z * w), 2.0f * (x * z - y * w));
    const f3 c1 = mk(2.0f * (x * y - z * w), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z + x * w));
    const f3 c2 = mk(2.0f * (x * z + y * w), 2.0f * (y * z - x * w), 1.0f - 2.0f * (x * x + y * y));
    const float denom = fabsf(dot3(c0, a0) + dot3(c1, a1) + dot3(c2, a2)) + 1e-9f;
    const f3 omega = (cross3(c0, a0) + cross3(c1, a1) + cross3(c2, a2)) * (1.0f / denom);
    const float wl = len3(omega);
    if (wl < 1e-9f) break;
    // quatFromAxisAngle (math/gpu.ts:27-30)
    const f3 axis = omega * (1.0f / wl);
    const float half = wl * 0.5f;
    float sh, ch;
    sin_cos(half, &sh, &ch);
    const f4 m = quat_mul(f4{axis.x * sh, axis.y * sh, axis.z * sh, ch}, q);
    const float l = sqrtf(m.x * m.x + m.y * m.y + m.z * m.z + m.w * m.w);
    q = f4{m.x / l, m.y / l, m.z / l, m.w / l};
  }
  return q;
}

// ownerOf (pbd.ts:131-140): the particle's body, or max_bodies when the slot is
// not inside that body's [first, first + count).
__device__ __forceinline__ uint32_t owner_of(const PbdK& k, const Parts& pt, const BodyDev* bodies, uint32_t i) {
  const uint32_t b = pt.body[i];
  if (b >= k.max_bodies) return k.max_bodies;
  const BodyDev& bd = bodies[b];
  if (i < bd.first || i >= bd.first + bd.count) return k.max_bodies;
  return b;
}

__device__ __forceinline__ f3 mix3(f3 a, f3 b, float t) {
  return mk(a.x * (1.0f - t) + b.x * t, a.y * (1.0f - t) + b.y * t, a.z * (1.0f - t) + b.z * t);
}

// kick (pbd.ts:144-170)
__global__ void __launch_bounds__(kGroup) k_pbd_kick(PbdK k, Parts pt, const BodyDev* __restrict__ bodies) {
  const uint32_t i = blockIdx.x * kGroup + threadIdx.x;
  if (i >= k.capacity) return;
  const uint32_t b = owner_of(k, pt, bodies, i);
  if (b != k.push_body || b >= k.max_bodies) return;
  if (bodies[b].flags & CLAY_BODY_KINEMATIC) return;
  const f3 dv = mk(k.kick[0], k.kick[1], k.kick[2]);
  st(pt.v, i, ld(pt.v, i) + dv);
  st(pt.u, i, ld(pt.u, i) + dv);
  st(pt.q, i, ld(pt.q, i) - dv * k.h);
}

// predict (pbd.ts:172-223)
__global__ void __launch_bounds__(kGroup)
    k_pbd_predict(PbdK k, ColliderDev col, Parts pt, const BodyDev* __restrict__ bodies) {
  const uint32_t i = blockIdx.x * kGroup + threadIdx.x;
  if (i >= k.capacity) return;
  const uint32_t b = owner_of(k, pt, bodies, i);
  if (b >= k.max_bodies) return;
  const uint32_t flags = bodies[b].flags;
  if (flags & CLAY_BODY_KINEMATIC) return;
  f3 acc = (flags & CLAY_BODY_WEIGHTLESS) ? mk(0.0f, 0.0f, 0.0f) : mk(k.g[0], k.g[1], k.g[2]);
  if (b == k.push_body) acc = acc + mk(k.push[0], k.push[1], k.push[2]);
  const f3 pos = ld(pt.p, i), prev = ld(pt.q, i), vel = ld(pt.v, i), velp = ld(pt.u, i);
  const float third = 1.0f / 3.0f;
  const float h23 = 2.0f * k.h / 3.0f;
  const f3 vpred = (vel * 4.0f - velp) * third + acc * h23;
  f3 q = (pos * 4.0f - prev) * third + vpred * h23;
  f3 prev_out = pos;
  if (col.nleaf > 0) {
    float dw;
    const int leaf = collider_pick(col, b, q, &dw);
    if (dw < k.radius) {
      const f3 n = collider_normal(col, b, leaf, q);
      const float push = k.radius - dw;
      q = q + n * push;
      prev_out = pos + n * push;
      const f3 dp = q - pos;
      q = q - (dp - n * dot3(dp, n)) * k.friction;
    }
  }
  st(pt.q, i, prev_out);
  st(pt.p, i, q);
}

// match (pbd.ts:227-297): one workgroup of kReduce lanes per body.
__global__ void __launch_bounds__(kReduce) k_pbd_match(PbdK k, Parts pt, BodyDev* __restrict__ bodies) {
  __shared__ float sh[9][kReduce];
  const uint32_t b = blockIdx.x;
  const int li = threadIdx.x;
  const uint32_t first = bodies[b].first, count = bodies[b].count;
  auto reduce = [&](int slots) {
    __syncthreads();
    for (int s = kReduce / 2; s > 0; s >>= 1) {
      if (li < s)
        for (int c = 0; c < slots; c++) sh[c][li] = sh[c][li] + sh[c][li + s];
      __syncthreads();
    }
  };
  f3 sp = mk(0.0f, 0.0f, 0.0f), sr = mk(0.0f, 0.0f, 0.0f);
  for (uint32_t t = (uint32_t)li; t < count; t += kReduce) {
    sp = sp + ld(pt.p, first + t);
    sr = sr + ld(pt.r, first + t);
  }
  sh[0][li] = sp.x;
  sh[1][li] = sp.y;
  sh[2][li] = sp.z;
  sh[3][li] = sr.x;
  sh[4][li] = sr.y;
  sh[5][li] = sr.z;
  reduce(6);
  const float fc = (float)count;
  const float inv = 1.0f / (fc < 1.0f ? 1.0f : fc);
  const f3 com = mk(sh[0][0], sh[1][0], sh[2][0]) * inv;
  const f3 rcom = mk(sh[3][0], sh[4][0], sh[5][0]) * inv;
  __syncthreads();
  f3 a0 = mk(0.0f, 0.0f, 0.0f), a1 = mk(0.0f, 0.0f, 0.0f), a2 = mk(0.0f, 0.0f, 0.0f);
  for (uint32_t t = (uint32_t)li; t < count; t += kReduce) {
    const f3 pr = ld(pt.p, first + t) - com;
    const f3 rr = ld(pt.r, first + t) - rcom;
    a0 = a0 + pr * rr.x;
    a1 = a1 + pr * rr.y;
    a2 = a2 + pr * rr.z;
  }
  sh[0][li] = a0.x;
  sh[1][li] = a0.y;
  sh[2][li] = a0.z;
  sh[3][li] = a1.x;
  sh[4][li] = a1.y;
  sh[5][li] = a1.z;
  sh[6][li] = a2.x;
  sh[7][li] = a2.y;
  sh[8][li] = a2.z;
  reduce(9);
  if (li == 0 && count > 0) {
    BodyDev& bd = bodies[b];
    bd.com[0] = com.x;
    bd.com[1] = com.y;
    bd.com[2] = com.z;
    bd.rest_com[0] = rcom.x;
    bd.rest_com[1] = rcom.y;
    bd.rest_com[2] = rcom.z;
    const f4 q = extract_rotation(mk(sh[0][0], sh[1][0], sh[2][0]), mk(sh[3][0], sh[4][0], sh[5][0]),
                                  mk(sh[6][0], sh[7][0], sh[8][0]), f4{bd.rot[0], bd.rot[1], bd.rot[2], bd.rot[3]},
                                  k.rot_iters);
    bd.rot[0] = q.x;
    bd.rot[1] = q.y;
    bd.rot[2] = q.z;
    bd.rot[3] = q.w;
  }
}

// apply (pbd.ts:299-340)
__global__ void __launch_bounds__(kGroup)
    k_pbd_apply(PbdK k, ColliderDev col, Parts pt, const BodyDev* __restrict__ bodies) {
  const uint32_t i = blockIdx.x * kGroup + threadIdx.x;
  if (i >= k.capacity) return;
  const uint32_t b = owner_of(k, pt, bodies, i);
  if (b >= k.max_bodies) return;
  const BodyDev& bd = bodies[b];
  if (bd.flags & CLAY_BODY_KINEMATIC) return;
  const f4 rot = f4{bd.rot[0], bd.rot[1], bd.rot[2], bd.rot[3]};
  const f3 com = mk(bd.com[0], bd.com[1], bd.com[2]);
  const f3 rcom = mk(bd.rest_com[0], bd.rest_com[1], bd.rest_com[2]);
  const f3 pos = ld(pt.p, i), rest = ld(pt.r, i);
  const f3 target = com + quat_rotate(rot, rest - rcom);
  f3 q = mix3(pos, target, fclamp_s(bd.stiffness, 0.0f, 1.0f));
  if (col.nleaf > 0) {
    float dw;
    const int leaf = collider_pick(col, b, q, &dw);
    if (dw < k.radius) {
      const f3 n = collider_normal(col, b, leaf, q);
      const float push = k.radius - dw;
      q = q + n * push;
      st(pt.q, i, ld(pt.q, i) + n * push);
    }
  }
  st(pt.p, i, q);
  if (bd.plasticity > 0.0f) {
    const f3 rnew = quat_rotate(quat_conj(rot), q - com) + rcom;
    st(pt.r, i, mix3(rest, rnew, fclamp_s(bd.plasticity, 0.0f, 1.0f)));
  }
}

// finalize (pbd.ts:342-373)
__global__ void __launch_bounds__(kGroup) k_pbd_finalize(PbdK k, Parts pt, const BodyDev* __restrict__ bodies) {
  const uint32_t i = blockIdx.x * kGroup + threadIdx.x;
  if (i >= k.capacity) return;
  const uint32_t b = owner_of(k, pt, bodies, i);
  if (b >= k.max_bodies) return;
  if (bodies[b].flags & CLAY_BODY_KINEMATIC) return;
  f3 v = (ld(pt.p, i) - ld(pt.q, i)) * (1.0f / fmax_s(k.h, 1e-6f)) * k.damping;
  const float speed = len3(v);
  if (speed > k.max_speed) v = v * (k.max_speed / speed);
  st(pt.u, i, ld(pt.v, i));
  st(pt.v, i, v);
}

}  // namespace

struct Solver {
  Ctx* ctx = nullptr;
  Field* world = nullptr;
  clay_solver_desc desc{};
  uint32_t capacity = 0, max_bodies = 0, declared = 0;
  float* planes = nullptr;  // 15 * capacity: pos, prev, vel, velPrev, rest (x, y, z each)
  uint32_t* body = nullptr;
  BodyDev* bodies = nullptr;
  float push[3] = {0, 0, 0}, kick[3] = {0, 0, 0};
  uint32_t push_body = 0xFFFFFFFFu;
  bool kicking = false;
};

namespace {

Parts parts_of(const Solver* s) {
  Parts pt{};
  float** dst[5] = {pt.p, pt.q, pt.v, pt.u, pt.r};
  for (int f = 0; f < 5; f++)
    for (int a = 0; a < 3; a++) dst[f][a] = s->planes + (size_t)(3 * f + a) * s->capacity;
  pt.body = s->body;
  return pt;
}

PbdK pbd_k(const Solver* s) {
  PbdK k{};
  for (int a = 0; a < 3; a++) {
    k.g[a] = (float)s->desc.gravity[a];
    k.push[a] = s->push[a];
    k.kick[a] = s->kick[a];
  }
  k.h = (float)(s->desc.dt / s->desc.substeps);  // this.dt / this.substeps, f64 then f32 (pbd.ts:109)
  k.radius = (float)s->desc.radius;
  k.friction = (float)s->desc.friction;
  k.damping = (float)s->desc.damping;
  k.max_speed = (float)s->desc.max_speed;
  k.push_body = s->push_body;
  k.capacity = s->capacity;
  k.max_bodies = s->max_bodies;
  k.rot_iters = s->desc.rotation_iterations;
  return k;
}

void solver_free(Solver* s) {
  if (!s) return;
  dfree(s->planes);
  dfree(s->body);
  dfree(s->bodies);
  delete s;
}

}  // namespace

}  // namespace clay

using namespace clay;

extern "C" void clay_solver_desc_default(clay_solver_desc* d) {
  if (!d) return;
  d->gravity[0] = 0.0;
  d->gravity[1] = -9.81;
  d->gravity[2] = 0.0;
  d->dt = 1.0 / 60.0;
  d->substeps = 2;
  d->rotation_iterations = 2;
  d->radius = 0.1;
  d->friction = 0.35;
  d->damping = 0.995;
  d->max_speed = 40.0;
}

extern "C" int clay_solver_create(clay_ctx* ctx, clay_field* world, const clay_solver_desc* desc,
                                  uint32_t capacity, uint32_t max_bodies, clay_solver** out) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c || !desc || !out) return fail(CLAY_ERR_INVALID, "clay_solver_create: NULL argument");
  if (capacity == 0 || max_bodies == 0) return fail(CLAY_ERR_INVALID, "ClaySolver: capacity and maxBodies must be positive");
  if (desc->substeps < 1) return fail(CLAY_ERR_INVALID, "ClaySolver: substeps must be at least 1");
  if (!(desc->dt > 0)) return fail(CLAY_ERR_INVALID, "ClaySolver: dt must be positive");
  if (desc->rotation_iterations < 0) return fail(CLAY_ERR_INVALID, "ClaySolver: rotationIterations must not be negative");
  Solver* s = new Solver();
  s->ctx = c;
  s->world = reinterpret_cast<Field*>(world);
  s->desc = *desc;
  s->capacity = capacity;
  s->max_bodies = max_bodies;
  int rc;
  if ((rc = dalloc(&s->planes, (size_t)15 * capacity)) || (rc = dalloc(&s->body, capacity)) ||
      (rc = dalloc(&s->bodies, max_bodies))) {
    solver_free(s);
    return rc;
  }
  hipStream_t st = c->stream;
  if ((rc = launch_fill_u32(c, st, reinterpret_cast<uint32_t*>(s->planes), (size_t)15 * capacity, 0u)) ||
      (rc = launch_fill_u32(c, st, s->body, capacity, 0xFFFFFFFFu)) ||
      (rc = launch_fill_u32(c, st, reinterpret_cast<uint32_t*>(s->bodies), sizeof(BodyDev) / 4 * max_bodies, 0u))) {
    solver_free(s);
    return rc;
  }
  if (hipStreamSynchronize(st) != hipSuccess) {
    solver_free(s);
    return fail(CLAY_ERR_HIP, "ClaySolver: initialisation failed");
  }
  *out = reinterpret_cast<clay_solver*>(s);
  return CLAY_OK;
}

extern "C" int clay_solver_destroy(clay_solver* sv) {
  Solver* s = reinterpret_cast<Solver*>(sv);
  if (s) (void)hipStreamSynchronize(s->ctx->stream);
  solver_free(s);
  return CLAY_OK;
}

extern "C" int clay_solver_set_bodies(clay_solver* sv, const clay_body_desc* b, uint32_t n) {
  Solver* s = reinterpret_cast<Solver*>(sv);
  if (!s || (n && !b)) return fail(CLAY_ERR_INVALID, "clay_solver_set_bodies: NULL argument");
  if (n > s->max_bodies) return fail(CLAY_ERR_INVALID, "ParticleSet: body limit " + std::to_string(s->max_bodies) + " reached");
  std::vector<BodyDev> h(n);
  for (uint32_t i = 0; i < n; i++) {
    BodyDev d{};
    d.first = b[i].first;
    d.count = b[i].count;
    d.flags = b[i].flags;
    d.stiffness = b[i].stiffness;
    d.plasticity = b[i].plasticity;
    d.rot[3] = 1.0f;  // addBody: rot = (0, 0, 0, 1), com = restCom = 0 (particles.ts:150-166)
    h[i] = d;
  }
  CLAY_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
  if (n) CLAY_HIP_CHECK(hipMemcpy(s->bodies, h.data(), sizeof(BodyDev) * n, hipMemcpyHostToDevice));
  s->declared = n;
  return CLAY_OK;
}

extern "C" int clay_solver_set_particles(clay_solver* sv, const float* pos, const float* rest, const uint32_t* body,
                                         uint32_t n) {
  Solver* s = reinterpret_cast<Solver*>(sv);
  if (!s || (n && (!pos || !rest || !body))) return fail(CLAY_ERR_INVALID, "clay_solver_set_particles: NULL argument");
  if (n > s->capacity) return fail(CLAY_ERR_INVALID, "clay_solver_set_particles: more particles than capacity");
  CLAY_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
  const size_t cap = s->capacity;
  for (int a = 0; a < 3; a++) {
    if (!n) break;
    CLAY_HIP_CHECK(hipMemcpy(s->planes + a * cap, pos + (size_t)a * n, sizeof(float) * n, hipMemcpyHostToDevice));
    CLAY_HIP_CHECK(hipMemcpy(s->planes + (3 + a) * cap, pos + (size_t)a * n, sizeof(float) * n,
                             hipMemcpyHostToDevice));
    CLAY_HIP_CHECK(hipMemset(s->planes + (6 + a) * cap, 0, sizeof(float) * n));
    CLAY_HIP_CHECK(hipMemset(s->planes + (9 + a) * cap, 0, sizeof(float) * n));
    CLAY_HIP_CHECK(hipMemcpy(s->planes + (12 + a) * cap, rest + (size_t)a * n, sizeof(float) * n,
                             hipMemcpyHostToDevice));
  }
  if (n) CLAY_HIP_CHECK(hipMemcpy(s->body, body, sizeof(uint32_t) * n, hipMemcpyHostToDevice));
  return CLAY_OK;
}

extern "C" int clay_solver_set_force(clay_solver* sv, uint32_t body, const double accel[3], const double impulse[3]) {
  Solver* s = reinterpret_cast<Solver*>(sv);
  if (!s || !accel || !impulse) return fail(CLAY_ERR_INVALID, "clay_solver_set_force: NULL argument");
  for (int a = 0; a < 3; a++) {
    s->push[a] = (float)accel[a];
    s->kick[a] = (float)impulse[a];
  }
  s->push_body = body;
  s->kicking = impulse[0] != 0 || impulse[1] != 0 || impulse[2] != 0;
  return CLAY_OK;
}

extern "C" int clay_solver_step(clay_solver* sv, clay_stream stream) {
  Solver* s = reinterpret_cast<Solver*>(sv);
  if (!s) return fail(CLAY_ERR_INVALID, "clay_solver_step: NULL solver");
  if (s->declared == 0) return CLAY_OK;
  Ctx* c = s->ctx;
  hipStream_t st = pick(c, stream);
  const PbdK k = pbd_k(s);
  const Parts pt = parts_of(s);
  ColliderDev none{};
  const ColliderDev& col = s->world ? s->world->col : none;
  const unsigned groups = (s->capacity + kGroup - 1) / kGroup;
  if (s->kicking) {
    LaunchTimer t(c, st, "pbd_kick");
    hipLaunchKernelGGL(k_pbd_kick, dim3(groups), dim3(kGroup), 0, st, k, pt, s->bodies);
    CLAY_HIP_CHECK(hipGetLastError());
    s->kicking = false;
  }
  for (int i = 0; i < s->desc.substeps; i++) {
    {
      LaunchTimer t(c, st, "pbd_predict");
      hipLaunchKernelGGL(k_pbd_predict, dim3(groups), dim3(kGroup), 0, st, k, col, pt, s->bodies);
    }
    {
      LaunchTimer t(c, st, "pbd_match");
      hipLaunchKernelGGL(k_pbd_match, dim3(s->declared), dim3(kReduce), 0, st, k, pt, s->bodies);
    }
    {
      LaunchTimer t(c, st, "pbd_apply");
      hipLaunchKernelGGL(k_pbd_apply, dim3(groups), dim3(kGroup), 0, st, k, col, pt, s->bodies);
    }
    {
      LaunchTimer t(c, st, "pbd_finalize");
      hipLaunchKernelGGL(k_pbd_finalize, dim3(groups), dim3(kGroup), 0, st, k, pt, s->bodies);
    }
    CLAY_HIP_CHECK(hipGetLastError());
  }
  return CLAY_OK;
}

extern "C" int clay_solver_get_particles(clay_solver* sv, int32_t what, float* host) {
  Solver* s = reinterpret_cast<Solver*>(sv);
  if (!s || !host) return fail(CLAY_ERR_INVALID, "clay_solver_get_particles: NULL argument");
  if (what < 0 || what > 4) return fail(CLAY_ERR_INVALID, "clay_solver_get_particles: unknown field");
  CLAY_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
  CLAY_HIP_CHECK(hipMemcpy(host, s->planes + (size_t)3 * what * s->capacity, sizeof(float) * 3 * s->capacity,
                           hipMemcpyDeviceToHost));
  return CLAY_OK;
}

extern "C" int clay_solver_get_bodies(clay_solver* sv, float* host) {
  Solver* s = reinterpret_cast<Solver*>(sv);
  if (!s || !host) return fail(CLAY_ERR_INVALID, "clay_solver_get_bodies: NULL argument");
  CLAY_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
  std::vector<BodyDev> h(s->max_bodies);
  CLAY_HIP_CHECK(hipMemcpy(h.data(), s->bodies, sizeof(BodyDev) * s->max_bodies, hipMemcpyDeviceToHost));
  for (uint32_t b = 0; b < s->max_bodies; b++) {
    float* o = host + 10 * (size_t)b;
    for (int a = 0; a < 3; a++) {
      o[a] = h[b].com[a];
      o[3 + a] = h[b].rest_com[a];
    }
    for (int a = 0; a < 4; a++) o[6 + a] = h[b].rot[a];
  }
  return CLAY_OK;
}

extern "C" int clay_solver_positions(clay_solver* sv, const float* out[3]) {
  Solver* s = reinterpret_cast<Solver*>(sv);
  if (!s || !out) return fail(CLAY_ERR_INVALID, "clay_solver_positions: NULL argument");
  for (int a = 0; a < 3; a++) out[a] = s->planes + (size_t)a * s->capacity;
  return CLAY_OK;
}
