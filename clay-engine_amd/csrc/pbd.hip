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
  float* aux = nullptr;     // 4 * capacity: restNormal (x, y, z), material
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
  dfree(s->aux);
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
  if ((rc = dalloc(&s->planes, (size_t)15 * capacity)) || (rc = dalloc(&s->aux, (size_t)4 * capacity)) ||
      (rc = dalloc(&s->body, capacity)) ||
      (rc = dalloc(&s->bodies, max_bodies))) {
    solver_free(s);
    return rc;
  }
  hipStream_t st = c->stream;
  if ((rc = launch_fill_u32(c, st, reinterpret_cast<uint32_t*>(s->planes), (size_t)15 * capacity, 0u)) ||
      (rc = launch_fill_u32(c, st, reinterpret_cast<uint32_t*>(s->aux), (size_t)4 * capacity, 0u)) ||
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

extern "C" int clay_solver_add_body(clay_solver* sv, const clay_body_desc* b, uint32_t* index) {
  Solver* s = reinterpret_cast<Solver*>(sv);
  if (!s || !b) return fail(CLAY_ERR_INVALID, "clay_solver_add_body: NULL argument");
  if (s->declared >= s->max_bodies)
    return fail(CLAY_ERR_INVALID, "ParticleSet: body limit " + std::to_string(s->max_bodies) + " reached");
  BodyDev d{};
  d.first = b->first;
  d.count = b->count;
  d.flags = b->flags;
  d.stiffness = b->stiffness;
  d.plasticity = b->plasticity;
  d.rot[3] = 1.0f;
  // Only the new record is written: the others keep the state the GPU maintains.
  CLAY_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
  CLAY_HIP_CHECK(hipMemcpy(s->bodies + s->declared, &d, sizeof(BodyDev), hipMemcpyHostToDevice));
  if (index) *index = s->declared;
  s->declared++;
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
  for (int a = 0; a < 4 && n; a++) CLAY_HIP_CHECK(hipMemset(s->aux + a * cap, 0, sizeof(float) * n));
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
  if (what < 0 || what > 6) return fail(CLAY_ERR_INVALID, "clay_solver_get_particles: unknown field");
  CLAY_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
  const size_t cap = s->capacity;
  const float* src = what < 5 ? s->planes + 3 * what * cap : s->aux + 3 * (what - 5) * cap;
  CLAY_HIP_CHECK(hipMemcpy(host, src, sizeof(float) * (what < 6 ? 3 : 1) * cap, hipMemcpyDeviceToHost));
  return CLAY_OK;
}

extern "C" int clay_solver_get_owners(clay_solver* sv, uint32_t* host) {
  Solver* s = reinterpret_cast<Solver*>(sv);
  if (!s || !host) return fail(CLAY_ERR_INVALID, "clay_solver_get_owners: NULL argument");
  CLAY_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
  CLAY_HIP_CHECK(hipMemcpy(host, s->body, sizeof(uint32_t) * s->capacity, hipMemcpyDeviceToHost));
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

extern "C" int clay_solver_get_slices(clay_solver* sv, uint32_t* host) {
  Solver* s = reinterpret_cast<Solver*>(sv);
  if (!s || !host) return fail(CLAY_ERR_INVALID, "clay_solver_get_slices: NULL argument");
  CLAY_HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
  std::vector<BodyDev> h(s->max_bodies);
  CLAY_HIP_CHECK(hipMemcpy(h.data(), s->bodies, sizeof(BodyDev) * s->max_bodies, hipMemcpyDeviceToHost));
  for (uint32_t b = 0; b < s->max_bodies; b++) {
    host[3 * (size_t)b] = h[b].first;
    host[3 * (size_t)b + 1] = h[b].count;
    host[3 * (size_t)b + 2] = h[b].flags;
  }
  return CLAY_OK;
}

extern "C" int clay_solver_positions(clay_solver* sv, const float* out[3]) {
  Solver* s = reinterpret_cast<Solver*>(sv);
  if (!s || !out) return fail(CLAY_ERR_INVALID, "clay_solver_positions: NULL argument");
  for (int a = 0; a < 3; a++) out[a] = s->planes + (size_t)a * s->capacity;
  return CLAY_OK;
}

// ---------------------------------------------------------------------------
// SurfaceExtractor (lib/src/sim/extract.ts:85-484): naive surface nets over a
// field's zero level set on a res^3 lattice, writing particles into a slice of the
// set and a quad index list.  The reference takes particle and quad slots with
// atomic counters, so their order is whatever the atomics give; here slots go in
// ascending flat cell index (quads: lattice point, then edge axis x, y, z), which is
// one of those orders and makes the extraction reproducible.  Each of the two grid
// passes is three launches: mark (per-cell work plus the rank within the
// workgroup), a scan of the workgroup counts, then the write.
// ---------------------------------------------------------------------------
namespace clay {

namespace {

constexpr int kExtGroup = 256;

struct ExtK {
  float origin[3];
  float cell;
  float vel[3];
  uint32_t preserve;
  uint32_t res, cells, body, base, capacity, quad_capacity, scene, set_capacity;
  const float* motion;  // ExtractMotion {center, pad, velocity, pad}, or null
};

// The 12 lattice-edge accumulation of one cell (extract.ts:156-164, 285-297).
struct Acc4 {
  float x, y, z, w;
};
__device__ __forceinline__ Acc4 crossing(f3 pa, float da, f3 pb, float db) {
  if (da * db >= 0.0f) return Acc4{0.0f, 0.0f, 0.0f, 0.0f};
  const float t = fclamp_s(da / (da - db), 0.0f, 1.0f);
  return Acc4{fmix(pa.x, pb.x, t), fmix(pa.y, pb.y, t), fmix(pa.z, pb.z, t), 1.0f};
}
__device__ __forceinline__ Acc4 operator+(Acc4 a, Acc4 b) { return Acc4{a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }

__device__ __forceinline__ f3 scan_origin(const ExtK& k) {
  f3 o = mk(k.origin[0], k.origin[1], k.origin[2]);
  if (k.motion) o = o + mk(k.motion[0], k.motion[1], k.motion[2]);
  return o;
}

__device__ __forceinline__ float field_dist(const ColliderDev& col, uint32_t scene, f3 p) {
  float d;
  (void)collider_pick(col, scene, p, &d);
  return d;
}

// Rank of this lane's `count` (0..3) among the workgroup's lanes in order, and the
// workgroup total, from two wave64 ballots per bit and a per-wave LDS table.
__device__ __forceinline__ uint32_t group_rank(uint32_t count, uint32_t* total) {
  __shared__ uint32_t wave_sum[kExtGroup / 64];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint64_t m0 = __ballot(count & 1u), m1 = __ballot((count >> 1) & 1u);
  const uint64_t below = lane ? (~0ull >> (64u - lane)) : 0ull;
  const uint32_t in_wave = (uint32_t)__popcll(m0 & below) + 2u * (uint32_t)__popcll(m1 & below);
  if (lane == 0) wave_sum[wave] = (uint32_t)__popcll(m0) + 2u * (uint32_t)__popcll(m1);
  __syncthreads();
  uint32_t before = 0, all = 0;
  for (uint32_t w = 0; w < kExtGroup / 64; w++) {
    if (w < wave) before += wave_sum[w];
    all += wave_sum[w];
  }
  *total = all;
  return before + in_wave;
}

__device__ __forceinline__ void lattice_of(const ExtK& k, uint32_t idx, uint32_t* i, uint32_t* j, uint32_t* l) {
  *i = idx % k.res;
  *j = (idx / k.res) % k.res;
  *l = idx / (k.res * k.res);
}

// reset (extract.ts:193-254): snapshot the body's centre and mean velocity into
// the motion record when asked to, before the cloud is replaced.
__global__ void __launch_bounds__(kReduce) k_ext_reset(ExtK k, Parts pt, const BodyDev* __restrict__ bodies, float* motion) {
  __shared__ float sh[6][kReduce];
  const int li = threadIdx.x;
  const uint32_t first = bodies[k.body].first, count = bodies[k.body].count;
  const bool keep = k.preserve != 0u;
  f3 ps = mk(0.0f, 0.0f, 0.0f), vs = mk(0.0f, 0.0f, 0.0f);
  if (keep)
    for (uint32_t t = (uint32_t)li; t < count && first + t < k.set_capacity; t += kReduce) {
      ps = ps + ld(pt.p, first + t);
      vs = vs + ld(pt.v, first + t);
    }
  sh[0][li] = ps.x; sh[1][li] = ps.y; sh[2][li] = ps.z;
  sh[3][li] = vs.x; sh[4][li] = vs.y; sh[5][li] = vs.z;
  __syncthreads();
  for (int s = kReduce / 2; s > 0; s >>= 1) {
    if (li < s)
      for (int c = 0; c < 6; c++) sh[c][li] = sh[c][li] + sh[c][li + s];
    __syncthreads();
  }
  if (li == 0 && keep && count > 0u) {
    const float inv = 1.0f / (float)count;
    for (int c = 0; c < 3; c++) {
      motion[c] = sh[c][0] * inv;
      motion[4 + c] = sh[3 + c][0] * inv;
    }
    motion[3] = 0.0f;
    motion[7] = 0.0f;
  }
}

// cell pass, mark (extract.ts:256-329): the cell's dual vertex, one Lipschitz snap,
// and its material; rank = 1 + rank among the workgroup's vertices, 0 for none.
__global__ void __launch_bounds__(kExtGroup)
    k_ext_cells(ExtK k, ColliderDev col, float* __restrict__ vtx, uint32_t* __restrict__ rank,
                uint32_t* __restrict__ group_count) {
  const uint32_t idx = blockIdx.x * kExtGroup + threadIdx.x;
  uint32_t has = 0;
  if (idx < k.cells) {
    uint32_t i, j, l;
    lattice_of(k, idx, &i, &j, &l);
    const float c = k.cell;
    const f3 p0 = scan_origin(k) + mk((float)i, (float)j, (float)l) * c;
    const f3 px = p0 + mk(c, 0.0f, 0.0f), py = p0 + mk(0.0f, c, 0.0f), pz = p0 + mk(0.0f, 0.0f, c);
    const f3 pxy = p0 + mk(c, c, 0.0f), pxz = p0 + mk(c, 0.0f, c), pyz = p0 + mk(0.0f, c, c);
    const f3 pxyz = p0 + mk(c, c, c);
    const uint32_t sc = k.scene;
    const float d0 = field_dist(col, sc, p0), dx = field_dist(col, sc, px), dy = field_dist(col, sc, py);
    const float dz = field_dist(col, sc, pz), dxy = field_dist(col, sc, pxy), dxz = field_dist(col, sc, pxz);
    const float dyz = field_dist(col, sc, pyz), dxyz = field_dist(col, sc, pxyz);
    const Acc4 acc = crossing(p0, d0, px, dx) + crossing(py, dy, pxy, dxy) + crossing(pz, dz, pxz, dxz) +
                     crossing(pyz, dyz, pxyz, dxyz) + crossing(p0, d0, py, dy) + crossing(px, dx, pxy, dxy) +
                     crossing(pz, dz, pyz, dyz) + crossing(pxz, dxz, pxyz, dxyz) + crossing(p0, d0, pz, dz) +
                     crossing(px, dx, pxz, dxz) + crossing(py, dy, pyz, dyz) + crossing(pxy, dxy, pxyz, dxyz);
    if (!(acc.w < 0.5f)) {
      has = 1;
      const float inv = 1.0f / acc.w;
      const f3 v = mk(acc.x * inv, acc.y * inv, acc.z * inv);
      float dv;
      const int lv = collider_pick(col, sc, v, &dv);
      const f3 n = collider_normal(col, sc, lv, v);
      const f3 p = v - n * dv;
      float dp;
      const int lp = collider_pick(col, sc, p, &dp);
      const float mat = collider_material(col, sc, lp, p);
      const size_t C = k.cells;
      vtx[idx] = p.x;
      vtx[C + idx] = p.y;
      vtx[2 * C + idx] = p.z;
      vtx[3 * C + idx] = n.x;
      vtx[4 * C + idx] = n.y;
      vtx[5 * C + idx] = n.z;
      vtx[6 * C + idx] = mat;
    }
  }
  uint32_t total;
  const uint32_t r = group_rank(has, &total);
  if (idx < k.cells) rank[idx] = has ? r + 1u : 0u;
  if (threadIdx.x == 0) group_count[blockIdx.x] = total;
}

// Exclusive scan of the workgroup counts (one workgroup); the total goes to *total.
__global__ void __launch_bounds__(1024) k_ext_scan(const uint32_t* __restrict__ cnt, uint32_t* __restrict__ off,
                                                    uint32_t n, uint32_t* __restrict__ total) {
  __shared__ uint32_t buf[1024];
  __shared__ uint32_t carry;
  const uint32_t t = threadIdx.x;
  if (t == 0) carry = 0;
  __syncthreads();
  for (uint32_t base = 0; base < n; base += 1024) {
    const uint32_t v = base + t < n ? cnt[base + t] : 0u;
    buf[t] = v;
    __syncthreads();
    for (uint32_t s = 1; s < 1024; s <<= 1) {
      const uint32_t add = t >= s ? buf[t - s] : 0u;
      __syncthreads();
      buf[t] += add;
      __syncthreads();
    }
    if (base + t < n) off[base + t] = carry + buf[t] - v;
    __syncthreads();
    if (t == 1023) carry += buf[1023];
    __syncthreads();
  }
  if (t == 0) *total = carry;
}

// cell pass, write: particle slot base + rank, or no vertex past the capacity.
__global__ void __launch_bounds__(kExtGroup)
    k_ext_emit(ExtK k, Parts pt, float* __restrict__ aux, uint32_t* __restrict__ owner, const float* __restrict__ vtx,
               const uint32_t* __restrict__ rank, const uint32_t* __restrict__ group_off, uint32_t* __restrict__ ids) {
  const uint32_t idx = blockIdx.x * kExtGroup + threadIdx.x;
  if (idx >= k.cells) return;
  const uint32_t r = rank[idx];
  if (r == 0u) {
    ids[idx] = 0u;
    return;
  }
  const uint32_t slot = group_off[blockIdx.x] + r - 1u;
  if (slot >= k.capacity) {
    ids[idx] = 0u;
    return;
  }
  const uint32_t id = k.base + slot;
  const size_t C = k.cells, S = k.set_capacity;
  const f3 p = mk(vtx[idx], vtx[C + idx], vtx[2 * C + idx]);
  const f3 vel = k.motion ? mk(k.motion[4], k.motion[5], k.motion[6]) : mk(k.vel[0], k.vel[1], k.vel[2]);
  st(pt.p, id, p);
  st(pt.q, id, p);
  st(pt.v, id, vel);
  st(pt.u, id, vel);
  st(pt.r, id, p);
  aux[id] = vtx[3 * C + idx];
  aux[S + id] = vtx[4 * C + idx];
  aux[2 * S + id] = vtx[5 * C + idx];
  aux[3 * S + id] = vtx[6 * C + idx];
  owner[id] = k.body;
  ids[idx] = id + 1u;
}

__device__ __forceinline__ uint32_t cell_id(const ExtK& k, const uint32_t* ids, uint32_t i, uint32_t j, uint32_t l) {
  return ids[i + k.res * (j + k.res * l)];
}

// quad pass, mark (extract.ts:353-404): which of the three edges leaving lattice
// point (i, j, l) towards +x, +y, +z straddle the surface with all four cells
// holding a vertex (bits 0-2), whether the low end is inside (bit 3), and the rank
// of its first quad within the workgroup (bits 4-).
__global__ void __launch_bounds__(kExtGroup)
    k_ext_quads(ExtK k, ColliderDev col, const uint32_t* __restrict__ ids, uint32_t* __restrict__ info,
                uint32_t* __restrict__ group_count) {
  const uint32_t idx = blockIdx.x * kExtGroup + threadIdx.x;
  uint32_t mask = 0, inside = 0;
  if (idx < k.cells) {
    uint32_t i, j, l;
    lattice_of(k, idx, &i, &j, &l);
    const float c = k.cell;
    const f3 o = scan_origin(k) + mk((float)i, (float)j, (float)l) * c;
    const uint32_t sc = k.scene;
    const float dl = field_dist(col, sc, o);
    inside = dl < 0.0f ? 1u : 0u;
    if (j >= 1 && l >= 1 && field_dist(col, sc, o + mk(c, 0.0f, 0.0f)) * dl < 0.0f &&
        cell_id(k, ids, i, j - 1, l - 1) && cell_id(k, ids, i, j, l - 1) && cell_id(k, ids, i, j, l) &&
        cell_id(k, ids, i, j - 1, l))
      mask |= 1u;
    if (i >= 1 && l >= 1 && field_dist(col, sc, o + mk(0.0f, c, 0.0f)) * dl < 0.0f &&
        cell_id(k, ids, i - 1, j, l - 1) && cell_id(k, ids, i - 1, j, l) && cell_id(k, ids, i, j, l) &&
        cell_id(k, ids, i, j, l - 1))
      mask |= 2u;
    if (i >= 1 && j >= 1 && field_dist(col, sc, o + mk(0.0f, 0.0f, c)) * dl < 0.0f &&
        cell_id(k, ids, i - 1, j - 1, l) && cell_id(k, ids, i, j - 1, l) && cell_id(k, ids, i, j, l) &&
        cell_id(k, ids, i - 1, j, l))
      mask |= 4u;
  }
  uint32_t total;
  const uint32_t n = (mask & 1u) + ((mask >> 1) & 1u) + ((mask >> 2) & 1u);
  const uint32_t r = group_rank(n, &total);
  if (idx < k.cells) info[idx] = mask | (inside << 3) | (r << 4);
  if (threadIdx.x == 0) group_count[blockIdx.x] = total;
}

// emit (extract.ts:331-351): two triangles per quad, wound by the low end's sign.
__device__ __forceinline__ void put_quad(const ExtK& k, uint32_t* indices, uint32_t q, uint32_t a, uint32_t b,
                                         uint32_t cc, uint32_t dd, bool flip) {
  if (q >= k.quad_capacity) return;
  uint32_t* o = indices + (size_t)q * 6;
  const uint32_t i0 = a - 1u, i1 = (flip ? dd : b) - 1u, i2 = cc - 1u, i3 = (flip ? b : dd) - 1u;
  o[0] = i0;
  o[1] = i1;
  o[2] = i2;
  o[3] = i0;
  o[4] = i2;
  o[5] = i3;
}

__global__ void __launch_bounds__(kExtGroup)
    k_ext_quads_write(ExtK k, const uint32_t* __restrict__ ids, const uint32_t* __restrict__ info,
                      const uint32_t* __restrict__ group_off, uint32_t* __restrict__ indices) {
  const uint32_t idx = blockIdx.x * kExtGroup + threadIdx.x;
  if (idx >= k.cells) return;
  const uint32_t f = info[idx];
  if ((f & 7u) == 0u) return;
  uint32_t i, j, l;
  lattice_of(k, idx, &i, &j, &l);
  const bool flip = ((f >> 3) & 1u) == 0u;  // emit(..., !inside)
  uint32_t q = group_off[blockIdx.x] + (f >> 4);
  if (f & 1u)
    put_quad(k, indices, q++, cell_id(k, ids, i, j - 1, l - 1), cell_id(k, ids, i, j, l - 1), cell_id(k, ids, i, j, l),
             cell_id(k, ids, i, j - 1, l), flip);
  if (f & 2u)
    put_quad(k, indices, q++, cell_id(k, ids, i - 1, j, l - 1), cell_id(k, ids, i - 1, j, l), cell_id(k, ids, i, j, l),
             cell_id(k, ids, i, j, l - 1), flip);
  if (f & 4u)
    put_quad(k, indices, q, cell_id(k, ids, i - 1, j - 1, l), cell_id(k, ids, i, j - 1, l), cell_id(k, ids, i, j, l),
             cell_id(k, ids, i - 1, j, l), flip);
}

// finish (extract.ts:406-441): the body's slice, centroid, identity rotation, and
// the indexed draw's arguments.
__global__ void __launch_bounds__(kReduce)
    k_ext_finish(ExtK k, Parts pt, BodyDev* __restrict__ bodies, const uint32_t* __restrict__ counters,
                 uint32_t* __restrict__ draw_args) {
  __shared__ float sh[3][kReduce];
  const int li = threadIdx.x;
  const uint32_t n = counters[0] < k.capacity ? counters[0] : k.capacity;
  f3 sum = mk(0.0f, 0.0f, 0.0f);
  for (uint32_t t = (uint32_t)li; t < n; t += kReduce) sum = sum + ld(pt.r, k.base + t);
  sh[0][li] = sum.x;
  sh[1][li] = sum.y;
  sh[2][li] = sum.z;
  __syncthreads();
  for (int s = kReduce / 2; s > 0; s >>= 1) {
    if (li < s)
      for (int c = 0; c < 3; c++) sh[c][li] = sh[c][li] + sh[c][li + s];
    __syncthreads();
  }
  if (li == 0) {
    const float inv = 1.0f / fmax_s((float)n, 1.0f);
    BodyDev& b = bodies[k.body];
    b.first = k.base;
    b.count = n;
    for (int c = 0; c < 3; c++) {
      b.rest_com[c] = sh[c][0] * inv;
      b.com[c] = sh[c][0] * inv;
    }
    b.rot[0] = 0.0f;
    b.rot[1] = 0.0f;
    b.rot[2] = 0.0f;
    b.rot[3] = 1.0f;
    const uint32_t q = counters[1] < k.quad_capacity ? counters[1] : k.quad_capacity;
    draw_args[0] = q * 6u;
    draw_args[1] = 1u;
    draw_args[2] = 0u;
    draw_args[3] = 0u;
  }
}

}  // namespace

struct Extractor {
  Solver* set = nullptr;
  Field* field = nullptr;
  clay_extractor_desc desc{};
  uint32_t cells = 0, groups = 0, quad_capacity = 0;
  float origin[3] = {0, 0, 0}, cell = 0.1f, vel[3] = {0, 0, 0};
  uint32_t preserve = 0;
  float* motion = nullptr;     // 8 floats, or null
  float* vtx = nullptr;        // 7 * cells: vertex xyz, normal xyz, material
  uint32_t* rank = nullptr;    // per cell, then per lattice point (quad info)
  uint32_t* ids = nullptr;     // per cell: particle id + 1, or 0
  uint32_t* gcount = nullptr;  // per workgroup
  uint32_t* goff = nullptr;    // per workgroup
  uint32_t* counters = nullptr;  // particles, quads (as the reference's counters read), draw args (4)
  uint32_t* indices = nullptr;   // quad_capacity * 6
};

namespace {

void extractor_free(Extractor* e) {
  if (!e) return;
  dfree(e->motion);
  dfree(e->vtx);
  dfree(e->rank);
  dfree(e->ids);
  dfree(e->gcount);
  dfree(e->goff);
  dfree(e->counters);
  dfree(e->indices);
  delete e;
}

ExtK ext_k(const Extractor* e) {
  ExtK k{};
  for (int a = 0; a < 3; a++) {
    k.origin[a] = e->origin[a];
    k.vel[a] = e->vel[a];
  }
  k.cell = e->cell;
  k.preserve = e->preserve;
  k.res = e->desc.resolution;
  k.cells = e->cells;
  k.body = e->desc.body;
  k.base = e->desc.base;
  k.capacity = e->desc.capacity;
  k.quad_capacity = e->quad_capacity;
  k.scene = e->desc.instance;
  k.set_capacity = e->set->capacity;
  k.motion = e->motion;
  return k;
}

}  // namespace

}  // namespace clay

extern "C" int clay_extractor_create(clay_solver* sv, clay_field* field, const clay_extractor_desc* d,
                                     clay_extractor** out) {
  Solver* s = reinterpret_cast<Solver*>(sv);
  Field* f = reinterpret_cast<Field*>(field);
  if (!s || !f || !d || !out) return fail(CLAY_ERR_INVALID, "clay_extractor_create: NULL argument");
  const uint32_t res = d->resolution ? d->resolution : 32u;
  if (res > 256) return fail(CLAY_ERR_INVALID, "SurfaceExtractor: resolution " + std::to_string(res) + " exceeds 256");
  if ((uint64_t)d->base + d->capacity > s->capacity)
    return fail(CLAY_ERR_INVALID, "SurfaceExtractor: slice [" + std::to_string(d->base) + ", " +
                                      std::to_string((uint64_t)d->base + d->capacity) + ") exceeds set capacity");
  if (d->body >= s->max_bodies)
    return fail(CLAY_ERR_INVALID, "SurfaceExtractor: body " + std::to_string(d->body) + " is outside the set's " +
                                      std::to_string(s->max_bodies) + " bodies");
  Extractor* e = new (std::nothrow) Extractor();
  if (!e) return fail(CLAY_ERR_INVALID, "out of host memory");
  e->set = s;
  e->field = f;
  e->desc = *d;
  e->desc.resolution = res;
  e->cells = res * res * res;
  e->groups = (e->cells + kExtGroup - 1) / kExtGroup;
  // A closed surface-nets mesh has about as many quads as vertices (extract.ts:110-112).
  e->quad_capacity = (uint32_t)std::ceil((double)d->capacity * 1.5);
  Ctx* c = s->ctx;
  int rc;
  if ((d->motion && (rc = dalloc(&e->motion, 8))) || (rc = dalloc(&e->vtx, (size_t)7 * e->cells)) ||
      (rc = dalloc(&e->rank, e->cells)) || (rc = dalloc(&e->ids, e->cells)) || (rc = dalloc(&e->gcount, e->groups)) ||
      (rc = dalloc(&e->goff, e->groups)) || (rc = dalloc(&e->counters, 8)) ||
      (rc = dalloc(&e->indices, (size_t)6 * std::max(e->quad_capacity, 1u)))) {
    extractor_free(e);
    return rc;
  }
  hipStream_t st = c->stream;
  if ((e->motion && (rc = launch_fill_u32(c, st, reinterpret_cast<uint32_t*>(e->motion), 8, 0u))) ||
      (rc = launch_fill_u32(c, st, e->ids, e->cells, 0u)) || (rc = launch_fill_u32(c, st, e->counters, 8, 0u)) ||
      (rc = launch_fill_u32(c, st, e->indices, (size_t)6 * std::max(e->quad_capacity, 1u), 0u))) {
    extractor_free(e);
    return rc;
  }
  if (hipStreamSynchronize(st) != hipSuccess) {
    extractor_free(e);
    return fail(CLAY_ERR_HIP, "SurfaceExtractor: initialisation failed");
  }
  *out = reinterpret_cast<clay_extractor*>(e);
  return CLAY_OK;
}

extern "C" int clay_extractor_destroy(clay_extractor* eh) {
  Extractor* e = reinterpret_cast<Extractor*>(eh);
  if (e) (void)hipStreamSynchronize(e->set->ctx->stream);
  extractor_free(e);
  return CLAY_OK;
}

extern "C" int clay_extractor_set_region(clay_extractor* eh, const double origin[3], double size, const double vel[3],
                                         int32_t preserve_motion) {
  Extractor* e = reinterpret_cast<Extractor*>(eh);
  if (!e || !origin) return fail(CLAY_ERR_INVALID, "clay_extractor_set_region: NULL argument");
  for (int a = 0; a < 3; a++) {
    e->origin[a] = (float)origin[a];
    e->vel[a] = vel ? (float)vel[a] : 0.0f;
  }
  e->cell = (float)(size / e->desc.resolution);  // size / this.resolution in f64, stored as f32
  e->preserve = preserve_motion ? 1u : 0u;
  return CLAY_OK;
}

extern "C" int clay_extractor_extract(clay_extractor* eh, clay_stream stream) {
  Extractor* e = reinterpret_cast<Extractor*>(eh);
  if (!e) return fail(CLAY_ERR_INVALID, "clay_extractor_extract: NULL extractor");
  Solver* s = e->set;
  Ctx* c = s->ctx;
  hipStream_t st = pick(c, stream);
  const ExtK k = ext_k(e);
  const Parts pt = parts_of(s);
  const ColliderDev& col = e->field->col;
  if (e->motion) {
    LaunchTimer t(c, st, "extract_reset");
    hipLaunchKernelGGL(k_ext_reset, dim3(1), dim3(kReduce), 0, st, k, pt, s->bodies, e->motion);
  }
  {
    LaunchTimer t(c, st, "extract_cells");
    hipLaunchKernelGGL(k_ext_cells, dim3(e->groups), dim3(kExtGroup), 0, st, k, col, e->vtx, e->rank, e->gcount);
  }
  hipLaunchKernelGGL(k_ext_scan, dim3(1), dim3(1024), 0, st, e->gcount, e->goff, e->groups, e->counters);
  {
    LaunchTimer t(c, st, "extract_emit");
    hipLaunchKernelGGL(k_ext_emit, dim3(e->groups), dim3(kExtGroup), 0, st, k, pt, s->aux, s->body, e->vtx, e->rank,
                       e->goff, e->ids);
  }
  {
    LaunchTimer t(c, st, "extract_quads");
    hipLaunchKernelGGL(k_ext_quads, dim3(e->groups), dim3(kExtGroup), 0, st, k, col, e->ids, e->rank, e->gcount);
  }
  hipLaunchKernelGGL(k_ext_scan, dim3(1), dim3(1024), 0, st, e->gcount, e->goff, e->groups, e->counters + 1);
  hipLaunchKernelGGL(k_ext_quads_write, dim3(e->groups), dim3(kExtGroup), 0, st, k, e->ids, e->rank, e->goff,
                     e->indices);
  hipLaunchKernelGGL(k_ext_finish, dim3(1), dim3(kReduce), 0, st, k, pt, s->bodies, e->counters, e->counters + 4);
  CLAY_HIP_CHECK(hipGetLastError());
  return CLAY_OK;
}

extern "C" int clay_extractor_read(clay_extractor* eh, uint32_t counts[2], uint32_t draw_args[4], uint32_t* indices) {
  Extractor* e = reinterpret_cast<Extractor*>(eh);
  if (!e) return fail(CLAY_ERR_INVALID, "clay_extractor_read: NULL extractor");
  CLAY_HIP_CHECK(hipStreamSynchronize(e->set->ctx->stream));
  uint32_t h[8];
  CLAY_HIP_CHECK(hipMemcpy(h, e->counters, sizeof(h), hipMemcpyDeviceToHost));
  if (counts) {
    counts[0] = h[0];
    counts[1] = h[1];
  }
  if (draw_args)
    for (int i = 0; i < 4; i++) draw_args[i] = h[4 + i];
  if (indices && e->quad_capacity)
    CLAY_HIP_CHECK(hipMemcpy(indices, e->indices, sizeof(uint32_t) * 6 * e->quad_capacity, hipMemcpyDeviceToHost));
  return CLAY_OK;
}

extern "C" int clay_extractor_quad_capacity(const clay_extractor* eh, uint32_t* out) {
  const Extractor* e = reinterpret_cast<const Extractor*>(eh);
  if (!e || !out) return fail(CLAY_ERR_INVALID, "clay_extractor_quad_capacity: NULL argument");
  *out = e->quad_capacity;
  return CLAY_OK;
}

extern "C" int clay_extractor_motion(clay_extractor* eh, const float* in, float* out) {
  Extractor* e = reinterpret_cast<Extractor*>(eh);
  if (!e) return fail(CLAY_ERR_INVALID, "clay_extractor_motion: NULL extractor");
  if (!e->motion) return fail(CLAY_ERR_INVALID, "SurfaceExtractor: created without a motion record");
  CLAY_HIP_CHECK(hipStreamSynchronize(e->set->ctx->stream));
  if (in) CLAY_HIP_CHECK(hipMemcpy(e->motion, in, sizeof(float) * 8, hipMemcpyHostToDevice));
  if (out) CLAY_HIP_CHECK(hipMemcpy(out, e->motion, sizeof(float) * 8, hipMemcpyDeviceToHost));
  return CLAY_OK;
}
