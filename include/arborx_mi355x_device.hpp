// arborx_mi355x_device.hpp -- device-side building blocks of the MI355X
// BVH, shared by the library's gfx950 kernels and by header-only user code
// (callback traversals, include/ArborX.hpp under hipcc): geometry and
// Morton helpers, the pre-order node layout, the Q16 node format and the
// stackless rope walk.
//
// Every float expression here is meant to be compiled with
// -ffp-contract=off and correctly rounded division/sqrt, so it rounds
// exactly like the reference's CPU path (no FMA contraction, IEEE division
// and sqrt).
#pragma once

#include "arborx_mi355x.h"

#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>

#define AXB_HD __host__ __device__ __forceinline__
#define AXB_D __device__ __forceinline__

namespace axb
{

// ---------------------------------------------------------------- layout
// Pre-order node array of 16-byte slots (uint4).  Word 3 of a record's
// first slot tells the kind: >= 0 internal node (value = number of slots of
// its whole subtree, i.e. the rope is `slot + skip`), < 0 leaf (value = ~index).
//   FP32 internal : 2 slots {lo.x lo.y lo.z skip} {hi.x hi.y hi.z 0}
//   Q16 internal  : 1 slot  {qlo.x|qlo.y<<16, qlo.z|qhi.x<<16, qhi.y|qhi.z<<16, skip}
//   point leaf    : 1 slot  {x y z ~index}
//   box leaf      : 2 slots {lo.x lo.y lo.z ~index} {hi.x hi.y hi.z 0}
// The left child of an internal node is the next record; its right child is
// the rope of its left child (HappyTreeFriends.hpp:73-78).  An internal node
// whose subtree holds at most FLAT_MAX leaves is not stored: its leaves
// follow each other in the array (a walk reaching the subtree tests them
// all, so results and their order are unchanged).
constexpr int FLAT_MAX = AXB_FLAT_SUBTREE_MAX;
constexpr int GEOM_POINTS = 0;
constexpr int GEOM_BOXES = 1;
constexpr int NODES_FP32 = 0;
constexpr int NODES_Q16 = 1;

template <int G>
constexpr int leaf_slots()
{
  return G == GEOM_POINTS ? 1 : 2;
}
template <int F>
constexpr int internal_slots()
{
  return F == NODES_FP32 ? 2 : 1;
}

// Kernel error flags (ORed into a device word, checked by the host).
constexpr uint32_t KERR_SPIN = 1u;   // bounded wait exceeded
constexpr uint32_t KERR_STACK = 2u;  // traversal stack overflow
constexpr uint32_t KERR_FRAME = 4u;  // Q16 box outside its frame
constexpr uint32_t KERR_LAYOUT = 8u; // node array inconsistent (guarded access)

struct Box6
{
  float lo[3];
  float hi[3];
};

// ---------------------------------------------------------------- Morton
// ArborX_MortonCode.hpp:64-72 (expandBitsBy<2>, 32-bit).
AXB_HD uint32_t expand2_u32(uint32_t x)
{
  x &= 0x000003ffu;
  x = (x ^ (x << 16)) & 0xff0000ffu;
  x = (x ^ (x << 8)) & 0x0300f00fu;
  x = (x ^ (x << 4)) & 0x030c30c3u;
  x = (x ^ (x << 2)) & 0x09249249u;
  return x;
}
// ArborX_MortonCode.hpp:188-197 (expandBitsBy<2>, 64-bit).
AXB_HD uint64_t expand2_u64(uint64_t x)
{
  x &= 0x1fffffull;
  x = (x | x << 32) & 0x1f00000000ffffull;
  x = (x | x << 16) & 0x1f0000ff0000ffull;
  x = (x | x << 8) & 0x100f00f00f00f00full;
  x = (x | x << 4) & 0x10c30c30c30c30c3ull;
  x = (x | x << 2) & 0x1249249249249249ull;
  return x;
}
// Kokkos::clamp semantics.
AXB_HD float kclamp(float v, float lo, float hi)
{
  return (v < lo) ? lo : ((hi < v) ? hi : v);
}
AXB_HD float kmin(float a, float b) { return (b < a) ? b : a; }
AXB_HD float kmax(float a, float b) { return (a < b) ? b : a; }

// ArborX_TranslateAndScale.hpp:23-37 (one axis).
AXB_HD float translate_and_scale(float c, float a, float b)
{
  return a != b ? (c - a) / (b - a) : 0.f;
}
// ArborX_MortonCode.hpp:318-347, DIM=3: N = 2^21 in float.
AXB_HD uint64_t morton64(float px, float py, float pz)
{
  const float N = 2097152.f, Nm1 = 2097151.f;
  uint64_t r = 0;
  r += expand2_u64((uint64_t)kclamp(px * N, 0.f, Nm1)) << 2;
  r += expand2_u64((uint64_t)kclamp(py * N, 0.f, Nm1)) << 1;
  r += expand2_u64((uint64_t)kclamp(pz * N, 0.f, Nm1));
  return r;
}
// ArborX_MortonCode.hpp:291-316, DIM=3: N = 2^10 in float.
AXB_HD uint32_t morton32(float px, float py, float pz)
{
  const float N = 1024.f, Nm1 = 1023.f;
  uint32_t r = 0;
  r += expand2_u32((uint32_t)kclamp(px * N, 0.f, Nm1)) << 2;
  r += expand2_u32((uint32_t)kclamp(py * N, 0.f, Nm1)) << 1;
  r += expand2_u32((uint32_t)kclamp(pz * N, 0.f, Nm1));
  return r;
}

// ---------------------------------------------------------------- distance
// Squared distances summed in the reference's order (ArborX_Distance.hpp:
// 52-70: t = b - a; s += t*t starting from 0).  The sqrt is applied by the
// caller only where a distance value is needed.
AXB_HD float dist2_pp(float ax, float ay, float az, float bx, float by,
                      float bz)
{
  float tx = bx - ax, ty = by - ay, tz = bz - az;
  float s = tx * tx;
  s = s + ty * ty;
  s = s + tz * tz;
  return s;
}
// ClosestPoint.hpp:48-66 + Distance.hpp:72-80.
AXB_HD float closest(float p, float lo, float hi)
{
  return p < lo ? lo : (p > hi ? hi : p);
}
AXB_HD float dist2_pbox(float px, float py, float pz, const Box6 &b)
{
  return dist2_pp(px, py, pz, closest(px, b.lo[0], b.hi[0]),
                  closest(py, b.lo[1], b.hi[1]),
                  closest(pz, b.lo[2], b.hi[2]));
}
// The same squared distance to an internal node's box, clamping with
// min / max (two instructions per axis instead of two compares and two
// selects).  For a box with lo <= hi -- every node box -- both clamps give
// the same d^2 for all inputs, NaN included (min / max return the non-NaN
// operand, and d = c - p is NaN whenever p is; the sign of a zero does not
// change its square).  Leaf boxes keep the reference's comparisons, since
// user data may hold inverted boxes.
// min(max(q, lo), hi) for lo <= hi (every node box): one median-of-three
// instruction on the GPU.  A NaN q gives a NaN distance either way, since
// the distance is taken from q itself.
AXB_HD float clamp_box(float q, float lo, float hi)
{
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_fmed3f(q, lo, hi);
#else
  return fminf(fmaxf(q, lo), hi);
#endif
}
AXB_HD float dist2_pnode(float px, float py, float pz, const Box6 &b)
{
  return dist2_pp(px, py, pz, clamp_box(px, b.lo[0], b.hi[0]), clamp_box(py, b.lo[1], b.hi[1]),
                  clamp_box(pz, b.lo[2], b.hi[2]));
}
// ArborX_Intersects.hpp:51-65.
// (Bitwise ors: every comparison is evaluated, so a walk's loads of the box
// are issued together instead of one after the other behind branches.)
AXB_HD bool intersects_bb(const Box6 &a, const Box6 &b)
{
  return !((a.lo[0] > b.hi[0]) | (a.hi[0] < b.lo[0]) | (a.lo[1] > b.hi[1]) |
           (a.hi[1] < b.lo[1]) | (a.lo[2] > b.hi[2]) | (a.hi[2] < b.lo[2]));
}

AXB_HD float bits_f(uint32_t u)
{
  union
  {
    uint32_t u;
    float f;
  } v;
  v.u = u;
  return v.f;
}
AXB_HD uint32_t f_bits(float f)
{
  union
  {
    uint32_t u;
    float f;
  } v;
  v.f = f;
  return v.u;
}

// Correctly rounded float sqrt (IEEE sqrt under
// -fhip-fp32-correctly-rounded-divide-sqrt on the device, libm on the host).
AXB_HD float sqrt_rn(float x) { return __builtin_sqrtf(x); }

// Largest float s with RN(sqrt(s)) <= r.  sqrt is monotone and never lands
// exactly on the midpoint m between r and the next float (m has 25
// significant bits, sqrt of a float that rounds there would need m^2 to be a
// float), so RN(sqrt(s)) <= r  <=>  s < m^2, and m^2 is exact in double.
AXB_HD float sqrt_threshold(float r)
{
  if (!(r >= 0.f))
    return r < 0.f ? -1.f : r; // negative radius: nothing; NaN: nothing
  if (r == INFINITY)
    return INFINITY;
  float rn = bits_f(f_bits(r) + 1u); // next float above r (may be +inf)
  double m = 0.5 * ((double)r + (double)rn);
  double m2 = m * m;
  float t = (float)m2;
  if ((double)t >= m2)
    t = bits_f(f_bits(t) - 1u); // step below m^2 (t > 0 here)
  return t;
}

// ---------------------------------------------------------------- Q16
// Node box quantised on the 65535-step grid of a frame box F, one formula
// for both faces:
//   v(q) = fma(q, s, F.lo),   s = fma(F.hi - F.lo, C, 2^-149),
//   C = 0x1.00010ap-16
// (correctly rounded fmas, identical on the host and the device).  C
// exceeds (1 + 2^-21) / 65535, so s > (F.hi - F.lo) / 65535 however the
// difference and the product round, and v(65535) >= F.hi; the added
// smallest denormal keeps s > 0 where the product would underflow.  qlo is the
// largest q with v(q) <= box.lo and qhi the smallest with v(q) >= box.hi
// (both exist for a box inside its frame), so the decoded box always
// contains the exact one.  Decoding needs the frame's lower corner and
// step only.
constexpr float Q16_INV_UP = 0x1.00010ap-16f;

AXB_HD float q16_scale(float flo, float fhi)
{
  return __builtin_fmaf(fhi - flo, Q16_INV_UP, 0x1p-149f);
}
AXB_HD float q16_lo(float flo, float s, uint32_t q)
{
  return __builtin_fmaf((float)q, s, flo);
}
// 1 / s: the hardware's approximate reciprocal on the device, a division on
// the host.  It only seeds the searches below, whose result does not
// depend on the seed.
AXB_HD float approx_rcp(float s)
{
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(s);
#else
  return 1.0f / s;
#endif
}

// Largest q in [0, 65535] with v(q) <= v (0 when there is none).  v(q) is
// non-decreasing in q; the search starts from the window around the
// estimate (v - flo) / s (rs: about 1 / s) and widens to the full range when
// the window does not bracket the answer, so the result is the plain binary
// search's.
AXB_HD uint32_t q16_lo_index(float flo, float s, float rs, float v)
{
  const float est = (v - flo) * rs;
  const int q = est >= 0.f ? (est <= 65535.f ? (int)est : 65535) : 0; // NaN -> 0
  int a = q - 2 < 0 ? 0 : q - 2, b = q + 2 > 65535 ? 65535 : q + 2;
  if ((a == 0 || q16_lo(flo, s, (uint32_t)a) <= v) &&
      (b == 65535 || q16_lo(flo, s, (uint32_t)(b + 1)) > v))
  {
    // the answer is in [a, b] and {j : v(j) <= v} is a prefix of it
    int cnt = 0;
#pragma unroll
    for (int j = 1; j < 5; ++j)
      cnt += (a + j <= b && q16_lo(flo, s, (uint32_t)(a + j)) <= v);
    return (uint32_t)(a + cnt);
  }
  a = 0, b = 65535;
  while (a < b)
  {
    const int m = (a + b + 1) >> 1;
    if (q16_lo(flo, s, (uint32_t)m) <= v)
      a = m;
    else
      b = m - 1;
  }
  return (uint32_t)a;
}

// Smallest q in [0, 65535] with v(q) >= v (65535 when there is none).
// Same bracketing as q16_lo_index.
AXB_HD uint32_t q16_hi_index(float flo, float s, float rs, float v)
{
  const float est = (v - flo) * rs;
  const int q = est >= 0.f ? (est <= 65535.f ? (int)est : 65535) : 0; // NaN -> 0
  int a = q - 2 < 0 ? 0 : q - 2, b = q + 2 > 65535 ? 65535 : q + 2;
  if ((b == 65535 || q16_lo(flo, s, (uint32_t)b) >= v) &&
      (a == 0 || q16_lo(flo, s, (uint32_t)(a - 1)) < v))
  {
    // the answer is in [a, b] and {j : v(j) >= v} is a suffix of it
    int cnt = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j)
      cnt += (a + j < b && q16_lo(flo, s, (uint32_t)(a + j)) < v);
    return (uint32_t)(a + cnt);
  }
  a = 0, b = 65535;
  while (a < b)
  {
    const int m = (a + b) >> 1;
    if (q16_lo(flo, s, (uint32_t)m) >= v)
      b = m;
    else
      a = m + 1;
  }
  return (uint32_t)a;
}

AXB_HD void q16_encode_axis(float flo, float fhi, float vlo, float vhi,
                            uint32_t &qlo, uint32_t &qhi)
{
  float s = q16_scale(flo, fhi);
  if (!(s > 0.f))
  {
    qlo = 0, qhi = 65535;
    return;
  }
  const float rs = approx_rcp(s);
  qlo = q16_lo_index(flo, s, rs, vlo);
  qhi = q16_hi_index(flo, s, rs, vhi);
}

// Decodes a Q16 record (words x, y, z) against a frame (lower corner flo,
// step s).
AXB_HD void q16_decode(uint32_t w0, uint32_t w1, uint32_t w2, const float flo[3],
                       const float s[3], Box6 &out)
{
  out.lo[0] = q16_lo(flo[0], s[0], w0 & 0xffffu);
  out.lo[1] = q16_lo(flo[1], s[1], w0 >> 16);
  out.lo[2] = q16_lo(flo[2], s[2], w1 & 0xffffu);
  out.hi[0] = q16_lo(flo[0], s[0], w1 >> 16);
  out.hi[1] = q16_lo(flo[1], s[1], w2 & 0xffffu);
  out.hi[2] = q16_lo(flo[2], s[2], w2 >> 16);
}

// ---------------------------------------------------------------- XCDs
// MI355X dispatches consecutive workgroups round-robin over its 8 XCDs,
// each with its own L2.  Traversal kernels process predicates in Morton
// order, so the logical block index is remapped to give every XCD one
// contiguous range of blocks: the predicates in flight on an XCD then share
// tree nodes in that XCD's L2.  A bijection of [0, nblocks).
constexpr uint32_t NUM_XCDS = 8;
AXB_HD uint32_t xcd_block(uint32_t b, uint32_t nblocks)
{
  const uint32_t per = nblocks / NUM_XCDS, rem = nblocks % NUM_XCDS;
  const uint32_t x = b % NUM_XCDS, i = b / NUM_XCDS;
  return x * per + (x < rem ? x : rem) + i;
}

// ---------------------------------------------------------------- walk
// Predicate kinds of the traversals.
constexpr int PRED_SPHERES = 0;
constexpr int PRED_BOXES = 1;
constexpr int PRED_RADIUS = 2; // points + one radius
constexpr int PRED_RAYS = 3;   // rays {origin, unit direction}, box trees

// ORs an error flag into a device word (nullptr: ignored).
AXB_HD void raise_error(uint32_t *err, uint32_t flag)
{
  if (err)
    __hip_atomic_fetch_or(err, flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Box of the internal record s0/s1.  Q16 records decode against the frame
// of their parent when reached by descending (desc), else against the scene.
// TV: any tree view with `Box6 scene` and `float ss[3]`.
template <int F, class TV>
AXB_HD void node_box(const TV &tv, uint4 s0, uint4 s1, bool desc, const Box6 &par,
                    const float *ps, Box6 &b)
{
  if (F == NODES_FP32)
  {
    b.lo[0] = bits_f(s0.x), b.lo[1] = bits_f(s0.y), b.lo[2] = bits_f(s0.z);
    b.hi[0] = bits_f(s1.x), b.hi[1] = bits_f(s1.y), b.hi[2] = bits_f(s1.z);
  }
  else
  {
    // frame chosen by selects rather than a branch: lanes of a wavefront
    // disagree on `desc` all the time
    float flo[3], fs[3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
      flo[d] = desc ? par.lo[d] : tv.scene.lo[d];
      fs[d] = desc ? ps[d] : tv.ss[d];
    }
    q16_decode(s0.x, s0.y, s0.z, flo, fs, b);
  }
}

template <int F>
AXB_HD void set_frame(const Box6 &b, Box6 &par, float *ps)
{
  if (F == NODES_Q16)
  {
    par = b;
    ps[0] = q16_scale(b.lo[0], b.hi[0]);
    ps[1] = q16_scale(b.lo[1], b.hi[1]);
    ps[2] = q16_scale(b.lo[2], b.hi[2]);
  }
}

// ---------------------------------------------------------------- rays
// A ray: origin and unit direction (ArborX_Ray.hpp:28-66).
struct RayPred
{
  float o[3];
  float d[3];
};

// A direction component equal to zero is taken as +0: the slab rule below
// treats -0 as a ray leaving every slab that strictly contains its origin,
// which would make a hit depend on the boxes of the nodes above the leaf.
AXB_HD RayPred ray_pred(const float *p6)
{
  RayPred r;
#pragma unroll
  for (int a = 0; a < 3; ++a)
  {
    r.o[a] = p6[a];
    r.d[a] = p6[3 + a] == 0.f ? 0.f : p6[3 + a];
  }
  return r;
}

// Parameter interval [tmin, tmax] of the ray inside box b by slabs
// (ArborX_Ray.hpp:107-153): per axis the parameters (face - origin) /
// direction of the two faces, entry first; an axis the ray runs parallel to
// gives (-inf, +inf) when the origin lies in the slab and an empty interval
// otherwise (the sign of direction * (face - origin) decides), and nothing
// at all when the origin is on the slab's lower face.  Non-empty iff
// tmin <= tmax.
AXB_HD bool ray_interval(const RayPred &r, const Box6 &b, float &tmin, float &tmax)
{
  tmin = -INFINITY;
  tmax = INFINITY;
#pragma unroll
  for (int a = 0; a < 3; ++a)
  {
    const float lo_o = b.lo[a] - r.o[a], hi_o = b.hi[a] - r.o[a];
    const float da = r.d[a];
    float t0, t1;
    if (da == 0.f)
    {
      const bool on_face = lo_o == 0.f;
      t0 = on_face ? -INFINITY : ((f_bits(da * hi_o) >> 31) ? INFINITY : -INFINITY);
      t1 = on_face ? INFINITY : ((f_bits(da * lo_o) >> 31) ? INFINITY : -INFINITY);
    }
    else
    {
      const float ta = lo_o / da, tb = hi_o / da;
      t0 = da > 0.f ? ta : tb;
      t1 = da > 0.f ? tb : ta;
    }
    if (tmin < t0)
      tmin = t0;
    if (tmax > t1)
      tmax = t1;
  }
  return tmin <= tmax;
}

// intersects(ray, box): the box meets the ray's half-line (ArborX_Ray.hpp:
// 156-167).
AXB_HD bool ray_hits(const RayPred &r, const Box6 &b)
{
  float t0, t1;
  return ray_interval(r, b, t0, t1) && t1 >= 0.f;
}

// distance(ray, box): the first parameter t >= 0 at which the ray is in the
// box, +inf if it never is (ArborX_Ray.hpp:429-444); -0 is returned as +0.
// A box containing another gets a distance no larger (the slab parameters
// are monotone in the faces), so node distances bound leaf distances from
// below, Q16 boxes (rounded outwards) included.
AXB_HD float ray_distance(const RayPred &r, const Box6 &b)
{
  float t0, t1;
  if (!(ray_interval(r, b, t0, t1) && t1 >= 0.f))
    return INFINITY;
  const float t = kmax(t0, 0.f);
  return t == 0.f ? 0.f : t;
}

// Spatial predicate of one lane: a sphere (centre, sqrt threshold of its
// radius), a box or a ray.
struct SpatialPred
{
  float cx, cy, cz, thr;
  Box6 qb;
  RayPred ray;
};

AXB_HD SpatialPred sphere_pred(float x, float y, float z, float r)
{
  SpatialPred p;
  p.cx = x, p.cy = y, p.cz = z, p.thr = sqrt_threshold(r);
  p.qb = Box6{{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
  p.ray = RayPred{{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
  return p;
}
AXB_HD SpatialPred box_pred(const Box6 &b)
{
  SpatialPred p;
  p.cx = p.cy = p.cz = p.thr = 0.f;
  p.qb = b;
  p.ray = RayPred{{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
  return p;
}
// p6 = {origin, direction}
AXB_HD SpatialPred ray_spatial_pred(const float *p6)
{
  SpatialPred p;
  p.cx = p6[0], p.cy = p6[1], p.cz = p6[2], p.thr = 0.f;
  p.qb = Box6{{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
  p.ray = ray_pred(p6);
  return p;
}

// State a point walk carries from one record to the next: the Q16 frame of
// the next record -- lower corner and step of the box just decoded when the
// walk descends into it, else the scene's -- chosen by selects at the end
// of each step (lanes of a wavefront disagree on descending all the time).
// The scene's frame is kept in per-lane registers (slo, sst): a select can
// read one scalar register at most besides its condition, so with the
// scene's values in scalar registers every step would first copy six of
// them to vector registers.
struct WalkFrame
{
  float lo[3], s[3];
  float slo[3], sst[3];
};

// x in a vector register, opaque to the compiler (so that it stays there)
AXB_HD float lane_copy(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
  float y;
  asm volatile("v_mov_b32 %0, %1" : "=v"(y) : "v"(x));
  return y;
#else
  return x;
#endif
}

template <class TV>
AXB_HD void walk_frame_init(const TV &tv, WalkFrame &fr)
{
#pragma unroll
  for (int d = 0; d < 3; ++d)
  {
    fr.slo[d] = lane_copy(tv.scene.lo[d]);
    fr.sst[d] = lane_copy(tv.ss[d]);
    fr.lo[d] = fr.slo[d], fr.s[d] = fr.sst[d];
  }
}

// Box of the internal record s0/s1 of a point walk (Q16: against fr).
template <int F>
AXB_HD void walk_node_box(uint4 s0, uint4 s1, const WalkFrame &fr, Box6 &b)
{
  if (F == NODES_FP32)
  {
    b.lo[0] = bits_f(s0.x), b.lo[1] = bits_f(s0.y), b.lo[2] = bits_f(s0.z);
    b.hi[0] = bits_f(s1.x), b.hi[1] = bits_f(s1.y), b.hi[2] = bits_f(s1.z);
  }
  else
    q16_decode(s0.x, s0.y, s0.z, fr.lo, fr.s, b);
}

// The frame of the next record: b when the walk descends into it.
template <int F, class TV>
AXB_HD void walk_frame_next(const TV &, bool desc, const Box6 &b, WalkFrame &fr)
{
  if (F == NODES_Q16)
  {
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
      fr.lo[d] = desc ? b.lo[d] : fr.slo[d];
      fr.s[d] = desc ? q16_scale(b.lo[d], b.hi[d]) : fr.sst[d];
    }
  }
}

// Exact test of a leaf record (s0, and s1 for a box leaf) against a sphere
// or box predicate: Intersects.hpp:51-114 with the sphere decided by the
// squared-distance threshold.
template <int G, int PK>
AXB_HD bool leaf_hit(const SpatialPred &pr, uint4 s0, uint4 s1)
{
  static_assert(!(G == GEOM_POINTS && PK == PRED_RAYS), "rays are tested against boxes");
  const float x = bits_f(s0.x), y = bits_f(s0.y), z = bits_f(s0.z);
  if (G == GEOM_POINTS)
  {
    if (PK == PRED_BOXES)
      return !((x > pr.qb.hi[0]) | (x < pr.qb.lo[0]) | (y > pr.qb.hi[1]) |
               (y < pr.qb.lo[1]) | (z > pr.qb.hi[2]) | (z < pr.qb.lo[2]));
    return dist2_pp(pr.cx, pr.cy, pr.cz, x, y, z) <= pr.thr;
  }
  const Box6 lb = {{x, y, z}, {bits_f(s1.x), bits_f(s1.y), bits_f(s1.z)}};
  if (PK == PRED_RAYS)
    return ray_hits(pr.ray, lb);
  return PK == PRED_BOXES ? intersects_bb(pr.qb, lb)
                          : dist2_pbox(pr.cx, pr.cy, pr.cz, lb) <= pr.thr;
}

// (Host-callable as well, so the CPU tests run this very walk over the
// oracle's node arrays.)
// Stackless pre-order walk over slots [p, end) (TreeTraversal.hpp:97-119:
// leaf -> test, next record; internal -> test ? first child : rope).  Each
// hit calls sink.hit(value), which returns false to stop the walk (the
// reference's CallbackTreeTraversalControl::early_exit,
// Callbacks.hpp:96-118).  The leaf record at `skip` is not reported.  A Q16
// walk starts at the root or at a record whose frame is the scene box.
// TV: a tree view with `const uint4 *slots`, `Box6 scene`, `float ss[3]` and
// `uint32_t *err`.  PK: PRED_BOXES tests the box, PRED_RAYS the ray (box
// leaves only), otherwise the sphere.
// Point leaves: one record per step, leaf and internal node through the
// same straight-line code.  With a sphere, the point test d2(c, p) and the
// node test d2(c, clamp(c, box)) are one computation once the closest point
// is chosen (the point itself for a leaf); with a box, the point test is
// the box test against [p, p] (the same six comparisons).  So the lanes of
// a wavefront, which disagree on the record kind in most steps, share the
// decode and test instructions instead of running both paths.
template <int F, int PK, class TV, class Sink>
AXB_HD void rope_walk_points(const TV &tv, const SpatialPred &pr, uint32_t p, uint32_t end,
                             uint32_t skip, Sink &sink, uint32_t &v_int, uint32_t &v_leaf)
{
  constexpr uint32_t SI = internal_slots<F>();
  WalkFrame fr;
  walk_frame_init(tv, fr);
  // The loop's only exits are its condition (a zero rope, reported below,
  // ends it as the end of the array does) and a sink asking to stop.
  uint32_t step = 1;
  while (step != 0 && p < end)
  {
    const uint4 s0 = tv.slots[p];
    const uint4 s1 = F == NODES_FP32 ? tv.slots[p + 1] : s0;
    const int32_t w = (int32_t)s0.w;
    const bool leaf = w < 0;
    Box6 bx; // meaningless for a leaf record, and unused then
    walk_node_box<F>(s0, s1, fr, bx);
    bool pass;
    if (PK == PRED_BOXES)
    {
      const float x = bits_f(s0.x), y = bits_f(s0.y), z = bits_f(s0.z);
      const Box6 tb = {{leaf ? x : bx.lo[0], leaf ? y : bx.lo[1], leaf ? z : bx.lo[2]},
                       {leaf ? x : bx.hi[0], leaf ? y : bx.hi[1], leaf ? z : bx.hi[2]}};
      pass = intersects_bb(pr.qb, tb);
    }
    else
    {
      const float cx = leaf ? bits_f(s0.x) : clamp_box(pr.cx, bx.lo[0], bx.hi[0]);
      const float cy = leaf ? bits_f(s0.y) : clamp_box(pr.cy, bx.lo[1], bx.hi[1]);
      const float cz = leaf ? bits_f(s0.z) : clamp_box(pr.cz, bx.lo[2], bx.hi[2]);
      pass = dist2_pp(pr.cx, pr.cy, pr.cz, cx, cy, cz) <= pr.thr;
    }
    v_leaf += leaf;
    v_int += !leaf;
    if (leaf && pass && p != skip && !sink.hit(~w))
      return;
    walk_frame_next<F>(tv, !leaf && pass, bx, fr);
    step = leaf ? 1u : (pass ? SI : (uint32_t)w);
    p += step;
  }
  if (step == 0)
    raise_error(tv.err, KERR_LAYOUT);
}

// Box leaves: one record per step, leaf and internal node through the same
// straight-line code, as rope_walk_points does for points.  The box tested
// is the leaf's exact one (slots p and p + 1) or the internal record's
// decoded one (FP32: the same two slots), picked by selects, so the lanes
// of a wavefront, which disagree on the record kind in most steps, run one
// path.  Box and ray predicates test both kinds of box alike; a sphere
// tests a leaf box with the reference's comparisons (leaf_hit) and a node
// box with clamps, both computed.  The read of slot p + 1 after a Q16
// internal record stays inside the array's zeroed tail.
template <int F, int PK, bool SLOTS, class TV, class Sink>
AXB_HD void rope_walk_boxes(const TV &tv, const SpatialPred &pr, uint32_t p, uint32_t end,
                            uint32_t skip, Sink &sink, uint32_t &v_int, uint32_t &v_leaf)
{
  constexpr uint32_t SL = leaf_slots<GEOM_BOXES>(), SI = internal_slots<F>();
  WalkFrame fr;
  walk_frame_init(tv, fr);
  uint32_t step = 1; // 0 after a zero rope (reported below), which ends the loop
  while (step != 0 && p < end)
  {
    const uint4 s0 = tv.slots[p];
    const uint4 s1 = tv.slots[p + 1];
    const int32_t w = (int32_t)s0.w;
    const bool leaf = w < 0;
    Box6 bx;
    walk_node_box<F>(s0, s1, fr, bx);
    if (F == NODES_Q16)
    {
      bx.lo[0] = leaf ? bits_f(s0.x) : bx.lo[0];
      bx.lo[1] = leaf ? bits_f(s0.y) : bx.lo[1];
      bx.lo[2] = leaf ? bits_f(s0.z) : bx.lo[2];
      bx.hi[0] = leaf ? bits_f(s1.x) : bx.hi[0];
      bx.hi[1] = leaf ? bits_f(s1.y) : bx.hi[1];
      bx.hi[2] = leaf ? bits_f(s1.z) : bx.hi[2];
    }
    bool pass;
    if (PK == PRED_RAYS)
      pass = ray_hits(pr.ray, bx);
    else if (PK == PRED_BOXES)
      pass = intersects_bb(pr.qb, bx);
    else
      pass = leaf ? dist2_pbox(pr.cx, pr.cy, pr.cz, bx) <= pr.thr
                  : dist2_pnode(pr.cx, pr.cy, pr.cz, bx) <= pr.thr;
    v_leaf += leaf;
    v_int += !leaf;
    if (leaf && pass && p != skip && !sink.hit(SLOTS ? (int32_t)p : ~w))
      return;
    walk_frame_next<F>(tv, !leaf && pass, bx, fr);
    step = leaf ? SL : (pass ? SI : (uint32_t)w);
    p += step;
  }
  if (step == 0)
    raise_error(tv.err, KERR_LAYOUT);
}

// SLOTS: the sink receives the slot of each hit leaf record instead of its
// value (read back by sort_ray_row).
template <int F, int G, int PK, bool SLOTS = false, class TV, class Sink>
AXB_HD void rope_walk(const TV &tv, const SpatialPred &pr, uint32_t p, uint32_t end,
                     uint32_t skip, Sink &sink, uint32_t &v_int, uint32_t &v_leaf)
{
  static_assert(!(G == GEOM_POINTS && (PK == PRED_RAYS || SLOTS)),
                "rays and slot rows are walks over box leaves");
  if constexpr (G == GEOM_POINTS)
    rope_walk_points<F, PK>(tv, pr, p, end, skip, sink, v_int, v_leaf);
  else
    rope_walk_boxes<F, PK, SLOTS>(tv, pr, p, end, skip, sink, v_int, v_leaf);
}

// ------------------------------------------------------- ordered walks
// Geometry of an ordered (best-first) query other than a ray:
// OrderedSpatial<Sphere | Box | Point> (detail/ArborX_Predicates.hpp:104-126)
// orders the values by distance(geometry, value) (Distance.hpp:52-208).
// kind: PRED_SPHERES {centre, r}, PRED_BOXES {lo, hi}, PRED_POINTS {p}.
constexpr int PRED_POINTS = 5;
struct OrderedGeom
{
  int kind;
  float q[6];
};
AXB_HD OrderedGeom ordered_geom(int kind, const float *p)
{
  OrderedGeom g;
  g.kind = kind;
  const int m = kind == PRED_SPHERES ? 4 : (kind == PRED_BOXES ? 6 : 3);
#pragma unroll
  for (int j = 0; j < 6; ++j)
    g.q[j] = j < m ? p[j] : 0.f;
  return g;
}
// distance(geometry, box): point-box (closest point, Distance.hpp:72-80);
// box-box (gaps of the separated axes, Distance.hpp:159-195: an axis where
// the boxes overlap adds +0, which leaves the sum unchanged);
// sphere-box (max(distance(centre, box) - r, 0), Distance.hpp:196-208).
// A point value is the box [p, p]: distance(point q, [p, p]) is
// distance(q, p), distance(box, [p, p]) is distance(p, box) and
// distance(sphere, [p, p]) is distance(p, sphere) bit for bit (the squared
// terms only change sign inside), the reference's reverse dispatch
// (ReverseDispatch.hpp:32-46) for the last two.
AXB_HD float ordered_distance(const OrderedGeom &g, const Box6 &b)
{
  if (g.kind == PRED_BOXES)
  {
    float s = 0.f;
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
      const float t = kmax(kmax(g.q[d] - b.hi[d], b.lo[d] - g.q[3 + d]), 0.f);
      s = s + t * t;
    }
    return sqrt_rn(s);
  }
  const float dc = sqrt_rn(dist2_pbox(g.q[0], g.q[1], g.q[2], b));
  return g.kind == PRED_SPHERES ? kmax(dc - g.q[3], 0.f) : dc;
}
AXB_HD float ordered_distance(const RayPred &r, const Box6 &b) { return ray_distance(r, b); }

// ------------------------------------------------------- best-first heap
// Binary min-heap of (key, payload) pairs in private arrays, for the
// best-first walk below.
template <int CAP>
struct OrderedHeap
{
  uint64_t key[CAP];
  uint64_t pay[CAP];
  int size = 0;

  AXB_HD bool push(uint64_t k, uint64_t p)
  {
    if (size >= CAP)
      return false;
    int j = size++;
    while (j > 0)
    {
      const int up = (j - 1) >> 1;
      if (key[up] <= k)
        break;
      key[j] = key[up], pay[j] = pay[up];
      j = up;
    }
    key[j] = k, pay[j] = p;
    return true;
  }
  // Puts (k, p) at position j and moves it down to its place.
  AXB_HD void sift_down(int j, uint64_t k, uint64_t p)
  {
    while (true)
    {
      int c = 2 * j + 1;
      if (c >= size)
        break;
      if (c + 1 < size && key[c + 1] < key[c])
        ++c;
      if (k <= key[c])
        break;
      key[j] = key[c], pay[j] = pay[c];
      j = c;
    }
    key[j] = k, pay[j] = p;
  }
  AXB_HD void pop(uint64_t &k, uint64_t &p)
  {
    k = key[0], p = pay[0];
    --size;
    if (size > 0)
      sift_down(0, key[size], pay[size]);
  }
};

// Heap key of a walk entry: distance bits (distances are >= +0, so their
// bits order like the values), then 0 for a node and 2^31 | index for a
// leaf.  At equal distance every node is opened before any leaf is
// reported, and leaves come out by index.
constexpr uint32_t ORDERED_LEAF = 0x80000000u;
AXB_HD uint64_t ordered_key(float d, uint32_t tie)
{
  return ((uint64_t)f_bits(d) << 32) | tie;
}

// Best-first walk for one ordered predicate, the reference's ordered
// spatial traversal (TreeTraversal.hpp:338-489): a ray (RayPred, trees of
// boxes: the leaves it hits) or the geometry of OrderedSpatial<Sphere | Box
// | Point> (OrderedGeom, any tree: every leaf, since their distances are
// finite); leaves are reported in increasing (distance, index) order,
// distance = ordered_distance (G: the trees' leaf geometry);
// sink.hit(value) returns false to stop (early exit,
// Callbacks.hpp:96-118).  The reference reports equal distances in the
// order of its heap; here they come by index.  Node distances bound those
// of the leaves below, so taking entries smallest key first reports leaves
// sorted.  An expanded node's children go to the heap except the smallest,
// which is taken next unless the heap holds a smaller key (the reference's
// closer-child step).  A node's children are its left child, the record
// after it, and the right child: the record at the left child's rope, or,
// when the left child is a run of leaves (a subtree that is not stored), the
// first internal record after the run inside the node's subtree.
// Q16: a left child decodes against its parent's box, so a node entry keeps
// (slot, slot of the first record of its left spine), that first record
// decoding against the scene, and is decoded down that spine when taken
// from the heap.  Returns false when the heap is full (KERR_STACK; the walk
// stops).  TV: a tree view with `slots`, `total`, `scene`, `ss`, `err`.
template <int F, int CAP = 64, int G = GEOM_BOXES, class TV, class P, class Sink>
AXB_HD bool ordered_walk(const TV &tv, const P &ray, Sink &sink, uint32_t &v_int,
                         uint32_t &v_leaf)
{
  constexpr uint32_t SL = leaf_slots<G>(), SI = internal_slots<F>();
  // a point leaf is the box [p, p]
  auto leaf_box = [](uint4 a, uint4 b) {
    return G == GEOM_POINTS
               ? Box6{{bits_f(a.x), bits_f(a.y), bits_f(a.z)}, {bits_f(a.x), bits_f(a.y), bits_f(a.z)}}
               : Box6{{bits_f(a.x), bits_f(a.y), bits_f(a.z)}, {bits_f(b.x), bits_f(b.y), bits_f(b.z)}};
  };
  OrderedHeap<CAP> h;
  const uint4 r0 = tv.slots[0], r1 = tv.slots[1];
  if ((int32_t)r0.w < 0)
  {
    // no stored node: at most FLAT_MAX leaves, followed by zeroed slots
    uint32_t q = 0;
    for (int j = 0; j < FLAT_MAX; ++j, q += SL)
    {
      const uint4 a0 = tv.slots[q];
      if ((int32_t)a0.w >= 0)
        break;
      ++v_leaf;
      const float d = ordered_distance(ray, leaf_box(a0, SL == 2 ? tv.slots[q + 1] : a0));
      if (d < INFINITY)
        h.push(ordered_key(d, ORDERED_LEAF | ~(uint32_t)a0.w), 0);
    }
    while (h.size > 0)
    {
      uint64_t k, pl;
      h.pop(k, pl);
      if (!sink.hit((int32_t)((uint32_t)k & ~ORDERED_LEAF)))
        break;
    }
    return true;
  }
  // current entry; for a node, cb is its box when have_box
  uint64_t ck = 0, cp = 0;
  Box6 cb;
  node_box<F>(tv, r0, r1, false, tv.scene, tv.ss, cb);
  ++v_int;
  if (!(ordered_distance(ray, cb) < INFINITY))
    return true;
  bool have_box = true;
  while (true)
  {
    if ((uint32_t)ck & ORDERED_LEAF)
    {
      if (!sink.hit((int32_t)((uint32_t)ck & ~ORDERED_LEAF)) || h.size == 0)
        return true;
      h.pop(ck, cp);
      have_box = false;
      continue;
    }
    const uint32_t p = (uint32_t)cp, anchor = (uint32_t)(cp >> 32);
    if (!have_box)
    {
      if (F == NODES_FP32)
        node_box<F>(tv, tv.slots[p], tv.slots[p + 1], false, tv.scene, tv.ss, cb);
      else
      {
        // down the left spine from the anchor (consecutive records)
        Box6 f = tv.scene;
        float fs[3] = {tv.ss[0], tv.ss[1], tv.ss[2]};
        for (uint32_t s = anchor;; ++s)
        {
          const uint4 q0 = tv.slots[s];
          node_box<F>(tv, q0, q0, s != anchor, f, fs, cb);
          if (s >= p)
            break;
          set_frame<F>(cb, f, fs);
        }
      }
    }
    Box6 frame = cb;
    float fs[3] = {0.f, 0.f, 0.f};
    set_frame<F>(cb, frame, fs);
    const uint32_t sub_end = p + (uint32_t)tv.slots[p].w;
    // the best child so far (bk, bp, bb) and the others pushed
    bool have_best = false, best_box = false;
    uint64_t bk = 0, bp = 0;
    Box6 bb;
    bool full = false;
    auto offer = [&](uint64_t k, uint64_t pl, bool is_node, const Box6 &b) {
      if (!have_best || k < bk)
      {
        if (have_best)
          full |= !h.push(bk, bp);
        have_best = true, bk = k, bp = pl, best_box = is_node, bb = b;
      }
      else
        full |= !h.push(k, pl);
    };
    auto leaf_child = [&](uint32_t q, uint4 a0) {
      ++v_leaf;
      const Box6 lb = leaf_box(a0, SL == 2 ? tv.slots[q + 1] : a0);
      const float d = ordered_distance(ray, lb);
      if (d < INFINITY)
        offer(ordered_key(d, ORDERED_LEAF | ~(uint32_t)a0.w), 0, false, lb);
    };
    auto node_child = [&](uint32_t q, uint4 a0, bool left) {
      ++v_int;
      Box6 nb;
      node_box<F>(tv, a0, F == NODES_FP32 ? tv.slots[q + 1] : a0, left, frame, fs, nb);
      const float d = ordered_distance(ray, nb);
      if (d < INFINITY)
        offer(ordered_key(d, 0u), (uint64_t)q | ((uint64_t)(left ? anchor : q) << 32), true, nb);
    };
    const uint32_t lp = p + SI;
    const uint4 l0 = tv.slots[lp];
    uint32_t q = lp;
    if ((int32_t)l0.w < 0)
    {
      // leaves of the left child, and of the right one too when it is not
      // stored either (at most 2 * FLAT_MAX leaves)
      for (int j = 0; j < 2 * FLAT_MAX && q < sub_end; ++j)
      {
        const uint4 a0 = tv.slots[q];
        if ((int32_t)a0.w >= 0)
          break;
        leaf_child(q, a0);
        q += SL;
      }
      if (q < sub_end)
      {
        const uint4 a0 = tv.slots[q];
        if ((int32_t)a0.w < 0)
        {
          raise_error(tv.err, KERR_LAYOUT);
          return false;
        }
        node_child(q, a0, false);
      }
    }
    else
    {
      node_child(lp, l0, true);
      q = lp + (uint32_t)l0.w;
      if (q >= sub_end)
      {
        raise_error(tv.err, KERR_LAYOUT);
        return false;
      }
      const uint4 a0 = tv.slots[q];
      if ((int32_t)a0.w >= 0)
        node_child(q, a0, false);
      else
      {
        // the right child's leaves fill the rest of the subtree
        for (int j = 0; j < FLAT_MAX && q < sub_end; ++j)
        {
          const uint4 b0 = tv.slots[q];
          if ((int32_t)b0.w >= 0)
            break;
          leaf_child(q, b0);
          q += SL;
        }
        if (q != sub_end)
        {
          raise_error(tv.err, KERR_LAYOUT);
          return false;
        }
      }
    }
    if (full)
    {
      raise_error(tv.err, KERR_STACK);
      return false;
    }
    if (!have_best)
    {
      if (h.size == 0)
        return true;
      h.pop(ck, cp);
      have_box = false;
    }
    else if (h.size > 0 && h.key[0] < bk)
    {
      // the heap's top comes first: it becomes current, the best child
      // takes its place
      ck = h.key[0], cp = h.pay[0];
      h.sift_down(0, bk, bp);
      have_box = false;
    }
    else
    {
      ck = bk, cp = bp, cb = bb;
      have_box = best_box;
    }
  }
}

// ------------------------------------------------------------ k nearest
AXB_HD bool pair_less(float da, int32_t ia, float db, int32_t ib)
{
  return da < db || (da == db && ia < ib);
}

// Binary max-heap on (distance, index) stored in the output row (global).
struct RowHeap
{
  int32_t *hi;
  float *hd;
  int k, count;
  float kth_d;
  int32_t kth_i;
  AXB_HD void init(int k_, int32_t *oi, float *od)
  {
    hi = oi, hd = od, k = k_, count = 0;
    kth_d = INFINITY, kth_i = INT_MAX;
  }
  AXB_HD bool full() const { return count >= k; }
  AXB_HD bool accepts(float dn, int32_t in) const
  {
    return count < k || pair_less(dn, in, kth_d, kth_i);
  }
  AXB_HD void insert(float dn, int32_t in)
  {
    if (count < k)
    {
      // push: sift up
      int j = count++;
      while (j > 0)
      {
        int par = (j - 1) >> 1;
        if (!pair_less(hd[par], hi[par], dn, in))
          break;
        hd[j] = hd[par], hi[j] = hi[par];
        j = par;
      }
      hd[j] = dn, hi[j] = in;
    }
    else
    {
      // pop-push: replace the top, sift down (PriorityQueue.hpp:87-93)
      int j = 0;
      while (true)
      {
        int l = 2 * j + 1;
        if (l >= k)
          break;
        int r = l + 1;
        int big = (r < k && pair_less(hd[l], hi[l], hd[r], hi[r])) ? r : l;
        if (!pair_less(dn, in, hd[big], hi[big]))
          break;
        hd[j] = hd[big], hi[j] = hi[big];
        j = big;
      }
      hd[j] = dn, hi[j] = in;
    }
    if (count >= k)
      kth_d = hd[0], kth_i = hi[0];
  }
  // heap sort ascending in place (Heap.hpp sortHeap)
  AXB_HD void finish(int per)
  {
    for (int end = per - 1; end > 0; --end)
    {
      float td = hd[0];
      int32_t ti = hi[0];
      float xd = hd[end];
      int32_t xi = hi[end];
      hd[end] = td, hi[end] = ti;
      int j = 0;
      while (true)
      {
        int l = 2 * j + 1;
        if (l >= end)
          break;
        int r = l + 1;
        int big = (r < end && pair_less(hd[l], hi[l], hd[r], hi[r])) ? r : l;
        if (!pair_less(xd, xi, hd[big], hi[big]))
          break;
        hd[j] = hd[big], hi[j] = hi[big];
        j = big;
      }
      hd[j] = xd, hi[j] = xi;
    }
  }
};

// (distance, index) packed into one 64-bit key: distances are >= 0, whose
// float bit patterns order like the values, so the key order is exactly the
// lexicographic (distance, index) order.
AXB_HD uint64_t pack_key(float d, int32_t i)
{
  return ((uint64_t)f_bits(d) << 32) | (uint32_t)i;
}

// Sorted list of K keys in registers; only the first k (k <= K) entries are
// meaningful.  Insertion is a branch-free network a'[j] = min(a[j],
// max(a[j-1], x)) with static indexing only.
template <int K>
struct RegList
{
  uint64_t a[K];
  int k, count;
  float kth_d;
  uint64_t kth;
  AXB_HD void init(int k_)
  {
    k = k_;
    count = 0;
#pragma unroll
    for (int j = 0; j < K; ++j)
      a[j] = ~0ull;
    kth_d = INFINITY;
    kth = ~0ull;
  }
  AXB_HD bool full() const { return count >= k; }
  AXB_HD bool accepts(float dn, int32_t in) const { return pack_key(dn, in) < kth; }
  AXB_HD void insert(float dn, int32_t in)
  {
    const uint64_t x = pack_key(dn, in);
#pragma unroll
    for (int j = K - 1; j > 0; --j)
    {
      const uint64_t hi = a[j - 1] < x ? x : a[j - 1];
      a[j] = hi < a[j] ? hi : a[j];
    }
    a[0] = x < a[0] ? x : a[0];
    ++count;
    uint64_t kk = a[0];
#pragma unroll
    for (int j = 1; j < K; ++j)
      if (j == k - 1)
        kk = a[j];
    if (count >= k)
    {
      kth = kk;
      kth_d = bits_f((uint32_t)(kk >> 32));
    }
  }
  AXB_HD void write(int32_t *oi, float *od, int per) const
  {
#pragma unroll
    for (int j = 0; j < K; ++j)
      if (j < per)
      {
        oi[j] = (int32_t)(uint32_t)a[j];
        if (od)
          od[j] = bits_f((uint32_t)(a[j] >> 32));
      }
  }
};

// Compare-exchange of two keys into ascending order (64-bit keys: compare
// and selects; 32-bit keys: v_min / v_max).
template <class T>
AXB_HD void cas_keys(T &x, T &y)
{
  const bool lt = x < y;
  const T lo = lt ? x : y;
  y = lt ? y : x;
  x = lo;
}

// Bitonic sorting network, ascending (N a power of two; static indices only,
// so the array stays in registers).
template <int N, class T>
AXB_HD void bitonic_sort(T (&a)[N])
{
#pragma unroll
  for (int size = 2; size <= N; size <<= 1)
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1)
#pragma unroll
      for (int i = 0; i < N; ++i)
      {
        const int j = i ^ stride;
        if (j > i)
        {
          if ((i & size) == 0)
            cas_keys(a[i], a[j]);
          else
            cas_keys(a[j], a[i]);
        }
      }
}

// Half-cleaner cascade: sorts a bitonic sequence ascending.
template <int N, class T>
AXB_HD void bitonic_merge(T (&a)[N])
{
#pragma unroll
  for (int stride = N >> 1; stride > 0; stride >>= 1)
#pragma unroll
    for (int i = 0; i < N; ++i)
    {
      const int j = i ^ stride;
      if (j > i)
        cas_keys(a[i], a[j]);
    }
}

// A row of leaf slots (rope_walk with SLOTS, or every leaf) becomes the
// row of their values in increasing (distance, index) order for an
// ordered predicate -- the rows of an ordered query without a best-first
// walk.  Up to 64 entries: keys (distance bits, index) sorted in registers
// by a bitonic network; longer rows: heap sort in place, keys recomputed
// from the records.  G: the leaf geometry (a point leaf is [p, p]).
template <int G, class TV, class P>
AXB_HD uint64_t ordered_leaf_key(const TV &tv, const P &pred, uint32_t p)
{
  const uint4 s0 = tv.slots[p], s1 = G == GEOM_BOXES ? tv.slots[p + 1] : s0;
  const Box6 b = {{bits_f(s0.x), bits_f(s0.y), bits_f(s0.z)},
                  {bits_f(s1.x), bits_f(s1.y), bits_f(s1.z)}};
  return ordered_key(ordered_distance(pred, b), ~s0.w);
}
template <int N, int G = GEOM_BOXES, class TV, class P>
AXB_HD void sort_ray_row_n(const TV &tv, const P &ray, int32_t *row, int m)
{
  uint64_t a[N];
#pragma unroll
  for (int j = 0; j < N; ++j)
    a[j] = j < m ? ordered_leaf_key<G>(tv, ray, (uint32_t)row[j]) : ~0ull;
  bitonic_sort<N>(a);
#pragma unroll
  for (int j = 0; j < N; ++j)
    if (j < m)
      row[j] = (int32_t)(uint32_t)a[j];
}
template <class TV>
AXB_HD uint64_t ray_leaf_key(const TV &tv, const RayPred &ray, int32_t slot)
{
  return ordered_leaf_key<GEOM_BOXES>(tv, ray, (uint32_t)slot);
}
template <int G = GEOM_BOXES, class TV, class P>
AXB_HD void sort_ray_row(const TV &tv, const P &ray, int32_t *row, int64_t m)
{
  if (m <= 8)
    return sort_ray_row_n<8, G>(tv, ray, row, (int)m);
  if (m <= 16)
    return sort_ray_row_n<16, G>(tv, ray, row, (int)m);
  if (m <= 32)
    return sort_ray_row_n<32, G>(tv, ray, row, (int)m);
  if (m <= 64)
    return sort_ray_row_n<64, G>(tv, ray, row, (int)m);
  // heap sort of the slots by key (max-heap, then the largest to the back)
  auto key = [&](int32_t slot) { return ordered_leaf_key<G>(tv, ray, (uint32_t)slot); };
  auto sift = [&](int64_t j, int64_t end) {
    const int32_t v = row[j];
    const uint64_t kv = key(v);
    while (true)
    {
      int64_t c = 2 * j + 1;
      if (c >= end)
        break;
      uint64_t kc = key(row[c]);
      if (c + 1 < end)
      {
        const uint64_t kr = key(row[c + 1]);
        if (kr > kc)
          ++c, kc = kr;
      }
      if (kv >= kc)
        break;
      row[j] = row[c];
      j = c;
    }
    row[j] = v;
  };
  for (int64_t j = m / 2 - 1; j >= 0; --j)
    sift(j, m);
  for (int64_t end = m - 1; end > 0; --end)
  {
    const int32_t t = row[0];
    row[0] = row[end];
    row[end] = t;
    sift(0, end);
  }
  for (int64_t j = 0; j < m; ++j)
    row[j] = ~(int32_t)tv.slots[(uint32_t)row[j]].w;
}


// Keys row[b, b + N) with entries from c on replaced by ~0.  The whole range
// lies inside the row's capacity (the caller's contract), so every load is
// issued unconditionally, 16 bytes at a time, all in flight together.
template <int N>
AXB_HD void load_keys(const uint2 *__restrict__ row, int b, int c, uint64_t (&a)[N])
{
  if constexpr (N % 2 == 0)
  {
    const uint4 *r4 = reinterpret_cast<const uint4 *>(row + b);
#pragma unroll
    for (int j = 0; j < N / 2; ++j)
    {
      const uint4 v = r4[j];
      a[2 * j] = b + 2 * j < c ? ((uint64_t)v.y << 32) | v.x : ~0ull;
      a[2 * j + 1] = b + 2 * j + 1 < c ? ((uint64_t)v.w << 32) | v.z : ~0ull;
    }
  }
  else
  {
#pragma unroll
    for (int j = 0; j < N; ++j)
    {
      const uint2 v = row[b + j];
      a[j] = b + j < c ? ((uint64_t)v.y << 32) | v.x : ~0ull;
    }
  }
}

// Key source over a collected row of {index, d2 bits} entries.  The row
// holds room for at least K entries and for c rounded up to a multiple of 8,
// and starts on a 16-byte boundary: entries past c are read and ignored.
struct RowKeys
{
  const uint2 *__restrict__ row;
  template <int N>
  AXB_HD void load(int b, int c, uint64_t (&a)[N]) const
  {
    load_keys<N>(row, b, c, a);
  }
  AXB_HD uint64_t at(int j) const
  {
    const uint2 v = row[j];
    return ((uint64_t)v.y << 32) | v.x;
  }
};

// Exact top-m (m <= K) of c candidates given as (d2 bits, index) keys by
// `src` (load<N>(b, c, a): keys b .. b+N-1, ~0 from c on; at(j): key j),
// written as (distance, index) ascending, distance = RN(sqrt(d2)), to oi
// (and od unless null).
//
// The K smallest keys are kept sorted in registers: the first K are sorted
// by a bitonic network, then each further batch of NB is sorted and merged
// in (the elementwise minimum of the ascending list and the reversed batch
// is a bitonic sequence holding the K smallest; a half-cleaner cascade
// sorts it).  About 13 compare-exchanges per candidate for K = 16, against
// a K-step insertion network per candidate.
//
// The reference orders by (distance, index) with distance = RN(sqrt(d2)).
// Sorting by (d2, index) gives the same order unless two different d2 round
// to the same distance among the outputs, or a candidate left out has the
// last output's distance with another d2; in that case (rare: the d2 are
// within a few ulps) the candidates are selected again with the distance
// keys themselves.
template <int K, class Src>
AXB_HD void select_keys(const Src &src, int c, int m, int32_t *__restrict__ oi,
                        float *__restrict__ od)
{
  constexpr int NB = K < 8 ? K : 8;
  uint64_t A[K];
  src.template load<K>(0, c, A);
  bitonic_sort<K>(A);
  uint64_t left = ~0ull; // smallest key dropped by a merge
  for (int b = K; b < c; b += NB)
  {
    uint64_t B[NB];
    src.template load<NB>(b, c, B);
    bitonic_sort<NB>(B);
#pragma unroll
    for (int j = K - NB; j < K; ++j)
    {
      const uint64_t x = A[j], y = B[K - 1 - j];
      const bool lt = x < y;
      A[j] = lt ? x : y;
      const uint64_t drop = lt ? y : x;
      left = drop < left ? drop : left;
    }
    bitonic_merge<K>(A);
  }
  // outputs written as the checks go, four at a time (one 16-byte store
  // each for a full row at 16-byte aligned addresses); a tie rewrites them
  // below
  bool tie = false;
  uint64_t last = A[0], next = left;
  float d_prev = 0.f, d_last = 0.f;
  const bool vec = K % 4 == 0 && m == K &&
                   (((uintptr_t)oi | (uintptr_t)(od ? od : reinterpret_cast<float *>(oi))) &
                    15u) == 0;
#pragma unroll
  for (int g = 0; g < K; g += 4)
  {
    int32_t ii[4] = {0, 0, 0, 0};
    float dd[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int u = 0; u < 4; ++u)
    {
      const int j = g + u;
      if (j < K && j < m)
      {
        const float dj = sqrt_rn(bits_f((uint32_t)(A[j] >> 32)));
        if (j > 0)
          tie = tie || (dj == d_prev && (uint32_t)(A[j] >> 32) != (uint32_t)(A[j - 1] >> 32));
        ii[u] = (int32_t)(uint32_t)A[j];
        dd[u] = dj;
        d_prev = dj;
        if (j == m - 1)
          last = A[j], d_last = dj;
      }
      if (j < K && j == m)
        next = A[j];
    }
    if (vec)
    {
      *reinterpret_cast<int4 *>(oi + g) = make_int4(ii[0], ii[1], ii[2], ii[3]);
      if (od)
        *reinterpret_cast<float4 *>(od + g) = make_float4(dd[0], dd[1], dd[2], dd[3]);
    }
    else
    {
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (g + u < K && g + u < m)
        {
          oi[g + u] = ii[u];
          if (od)
            od[g + u] = dd[u];
        }
    }
  }
  // a left-out candidate with the last distance but another d2 (an exact
  // repeat of the last d2 is ordered by index already; one behind it may
  // not be)
  const uint32_t h_last = (uint32_t)(last >> 32);
  if (!tie && next != ~0ull && sqrt_rn(bits_f((uint32_t)(next >> 32))) == d_last)
  {
    if ((uint32_t)(next >> 32) != h_last)
      tie = true;
    else
      for (int j = 0; j < c && !tie; ++j)
      {
        const uint32_t h = (uint32_t)(src.at(j) >> 32);
        tie = h != h_last && sqrt_rn(bits_f(h)) == d_last;
      }
  }
  if (!tie)
    return;
  RegList<K> L;
  L.init(m);
  for (int j = 0; j < c; ++j)
  {
    const uint64_t e = src.at(j);
    const float dist = sqrt_rn(bits_f((uint32_t)(e >> 32)));
    const int32_t idx = (int32_t)(uint32_t)e;
    if (L.accepts(dist, idx))
      L.insert(dist, idx);
  }
  L.write(oi, od, m);
}

// select_keys over the row at cand + row_at, written to oi_base + out_at
// (and od_base + out_at unless null).  (Base, offset) pairs rather than
// pointers: the bases stay scalar, which keeps the caller's walk within its
// register budget.
template <int K>
AXB_HD void select_row(const uint2 *__restrict__ cand, uint32_t row_at, int c, int m,
                       int32_t *__restrict__ oi_base, float *__restrict__ od_base,
                       int64_t out_at)
{
  const RowKeys src{cand + row_at};
  select_keys<K>(src, c, m, oi_base + out_at, od_base ? od_base + out_at : nullptr);
}

} // namespace axb
