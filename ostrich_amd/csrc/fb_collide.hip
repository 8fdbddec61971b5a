// fb_collide.hip -- primitive contact generation for the benchmark scenes on the GPU:
// the step before the solver (Newton's collide in the reference, base_simulator.py:118),
// producing Newton-format flat contacts (normal from shape0 to shape1, points in the frame
// of their body, world frame for static shapes, zero margins).
//
// The candidate set of a world is fixed by the model: for each shape pair (a, b) with a
// on a body and b a plane or a box, the sample points of a (8 box corners; for a cylinder
// the lowest point of its middle circle against a plane, 3 x k rim points against a box;
// a sphere's centre), then one point (or four grid cells) per body sphere against a
// heightfield.  A candidate
// is a contact when its signed distance is below the gap (and, against a box, its
// lateral excess is within the slack on the non-normal axes).  Contacts are emitted in
// (world, pair, point) order, capped at W * max_per_world, in two passes: one workgroup per
// world counts its contacts, one exclusive scan over worlds, and a second pass writes each
// world's contacts at its offset.  Each pass first drops the pairs whose shapes cannot
// touch (bounding sphere of shape0 against the plane or the box grown by the thresholds,
// a conservative test, so no contact is lost) and compacts the candidate points of the
// remaining pairs in their original order.  Fixed-size outputs and no host synchronisation: the
// pass can be captured into a graph together with the solver step.

#include <hip/hip_runtime.h>

#include <vector>

#include "fb_math.h"
#include "fb_solver.h"

enum { CGEO_PLANE = 0, CGEO_BOX = 1, CGEO_CYLINDER = 2, CGEO_SPHERE = 3, CGEO_HEIGHTFIELD = 4 };

struct ofb_collider {
  ofb_collider_desc d;  // pointers below are device copies
  int32_t *pairs, *pair_off, *geo, *body, *flat_id, *hf_spheres;
  float *xform, *size, *hf_heights;
  int32_t* world_count;   // [W]
  int32_t* world_offset;  // [W + 1]
};

namespace {

using namespace ofb;

constexpr int kColThreads = 256;

struct Cand {
  bool valid;
  v3 p0, p1, n;  // body-frame points, Newton normal (shape0 -> shape1)
  int s0, s1;
};

__device__ __forceinline__ tf ld_pose(const float* p) { return tf{mk(p[0], p[1], p[2]), qt{p[3], p[4], p[5], p[6]}}; }
__device__ __forceinline__ v3 qrot_inv(qt q, v3 v) { return qrot(qt{-q.x, -q.y, -q.z, q.w}, v); }

struct ColCtx {
  ofb_collider_desc d;
  const int32_t *geo, *body, *flat_id;
  const float *xform, *size, *body_q;
  int w;
  __device__ int col(int c) const { return w * d.shape_cols + c; }
  // world pose of the body of shape column c (identity for static shapes)
  __device__ tf body_pose(int c) const {
    const int b = body[col(c)];
    return b < 0 ? tident() : ld_pose(body_q + ((size_t)w * d.bodies_per_world + b) * 7);
  }
  __device__ tf shape_pose(int c) const {
    const tf X = ld_pose(xform + (size_t)col(c) * 7);
    if (body[col(c)] < 0) return X;
    const tf B = body_pose(c);
    return tf{B.p + qrot(B.q, X.p), qmul(B.q, X.q)};
  }
  __device__ v3 sz(int c) const { const float* s = size + (size_t)col(c) * 3; return mk(s[0], s[1], s[2]); }
};

// sample point s of shape column a in its shape frame (against a box); ang_cs / ang_sn
// hold cos / sin of j * 2 pi / k for the cylinder rim
__device__ __forceinline__ v3 sample_point(const ColCtx& C, int a, int s, const float* ang_cs, const float* ang_sn) {
  const int g = C.geo[C.col(a)];
  const v3 h = C.sz(a);
  if (g == CGEO_BOX) {  // corners in (x, y, z) sign order -,-,- ... +,+,+
    return mk(((s >> 2) & 1) ? h.x : -h.x, ((s >> 1) & 1) ? h.y : -h.y, (s & 1) ? h.z : -h.z);
  }
  if (g == CGEO_CYLINDER) {  // 3 rings (y = -hh, 0, +hh) of k rim points, axis = local y
    const int k = C.d.cylinder_samples;
    const int ring = s / k, j = s % k;
    return mk(ang_cs[j] * h.x, (float)(ring - 1) * h.y, ang_sn[j] * h.x);
  }
  return mk(0.f, 0.f, 0.f);  // sphere: centre
}

__device__ Cand eval_pair_point(const ColCtx& C, int a, int b, int s, const float* ang_cs, const float* ang_sn) {
  Cand o;
  o.valid = false;
  o.s0 = C.flat_id[C.col(a)];
  o.s1 = C.flat_id[C.col(b)];
  const int ga = C.geo[C.col(a)], gb = C.geo[C.col(b)];
  const tf Xa = C.shape_pose(a), Xb = C.shape_pose(b);
  const tf Ba = C.body_pose(a), Bb = C.body_pose(b);
  v3 pw, p1w, n_out;
  float dist;
  bool lat_ok = true;
  if (gb == CGEO_PLANE) {
    const v3 nrm = qrot(Xb.q, mk(0.f, 0.f, 1.f));
    if (ga == CGEO_CYLINDER) {  // lowest point of the middle circle
      const v3 ax = qrot(Xa.q, mk(0.f, 1.f, 0.f));
      v3 dd = (nrm - ax * dot(nrm, ax)) * -1.f;
      dd = dd * (1.f / fmaxf(sqrtf(dot(dd, dd)), 1e-9f));
      pw = Xa.p + dd * C.sz(a).x;
    } else if (ga == CGEO_SPHERE) {
      pw = Xa.p - nrm * C.sz(a).x;
    } else {
      pw = Xa.p + qrot(Xa.q, sample_point(C, a, s, ang_cs, ang_sn));
    }
    dist = dot(pw - Xb.p, nrm);
    p1w = pw - nrm * dist;
    n_out = nrm;
  } else {  // box b
    const v3 hb = C.sz(b);
    pw = Xa.p + qrot(Xa.q, sample_point(C, a, s, ang_cs, ang_sn));
    const v3 q = qrot_inv(Xb.q, pw - Xb.p);
    const float ex[3] = {fabsf(q.x) - hb.x, fabsf(q.y) - hb.y, fabsf(q.z) - hb.z};
    int axis = 0;
    if (C.body[C.col(b)] >= 0) {  // dynamic box below: face from the relative centre
      const v3 dc = qrot_inv(Xb.q, Xa.p - Xb.p);
      const v3 ha = C.sz(a);
      const float hn = sqrtf(dot(ha, ha));
      const float sc[3] = {fabsf(dc.x) / (hb.x + hn), fabsf(dc.y) / (hb.y + hn), fabsf(dc.z) / (hb.z + hn)};
      axis = sc[1] > sc[0] ? 1 : 0;
      if (sc[2] > sc[axis]) axis = 2;
    } else {
      axis = ex[1] > ex[0] ? 1 : 0;
      if (ex[2] > ex[axis]) axis = 2;
    }
    const float qa[3] = {q.x, q.y, q.z}, ha3[3] = {hb.x, hb.y, hb.z};
    float qp[3], nl[3];
    for (int k = 0; k < 3; ++k) {
      const float sg = qa[k] >= 0.f ? 1.f : -1.f;
      if (k == axis) {
        qp[k] = sg * ha3[k];
        nl[k] = sg;
      } else {
        qp[k] = fmaxf(fminf(qa[k], ha3[k]), -ha3[k]);
        nl[k] = 0.f;
        if (!(ex[k] <= C.d.lateral_slack)) lat_ok = false;
      }
    }
    dist = ex[axis];
    const v3 nw = qrot(Xb.q, mk(nl[0], nl[1], nl[2]));
    if (ga == CGEO_SPHERE) {
      const float r = C.sz(a).x;
      dist -= r;
      pw = pw - nw * r;
    }
    n_out = nw;
    p1w = Xb.p + qrot(Xb.q, mk(qp[0], qp[1], qp[2]));
  }
  o.valid = dist < C.d.gap && lat_ok;
  o.p0 = qrot_inv(Ba.q, pw - Ba.p);
  o.p1 = qrot_inv(Bb.q, p1w - Bb.p);
  o.n = n_out * -1.f;
  return o;
}

// closest point of triangle (a, b, c) to p and whether it lies in the face's interior
// region (Ericson, Real-Time Collision Detection, 5.1.5: the Voronoi region of p among the
// vertices, edges and face, tested in that order)
__device__ v3 closest_on_triangle(v3 p, v3 a, v3 b, v3 c, bool& face) {
  face = false;
  const v3 ab = b - a, ac = c - a, ap = p - a;
  const float d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.f && d2 <= 0.f) return a;
  const v3 bp = p - b;
  const float d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.f && d4 <= d3) return b;
  const float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f) return a + ab * (d1 / (d1 - d3));
  const v3 cp = p - c;
  const float d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.f && d5 <= d6) return c;
  const float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.f && d2 >= 0.f && d6 <= 0.f) return a + ac * (d2 / (d2 - d6));
  const float va = d3 * d6 - d5 * d4;
  if (va <= 0.f && d4 - d3 >= 0.f && d5 - d6 >= 0.f) return b + (c - b) * ((d4 - d3) / ((d4 - d3) + (d5 - d6)));
  face = true;
  const float denom = 1.f / (va + vb + vc);
  return a + ab * (vb * denom) + ac * (vc * denom);
}

// contact of a sphere centre c (radius r) with triangle (A, B, C) of the terrain: inside
// the face region the face's (signed distance along its upward normal: the centre may be
// below it); beside it, the nearest edge or vertex, for a centre in front of the face
// only (behind it, a neighbouring triangle holds the contact), along the face normal
// when the centre is within r/4 of that edge or vertex
struct TriContact {
  v3 foot, nrm;
  float dist;
  bool usable, face;
};
__device__ TriContact tri_contact(v3 c, float r, v3 A, v3 B, v3 C) {
  TriContact t;
  v3 nt = cross(B - A, C - A);
  nt = nt * (1.f / sqrtf(dot(nt, nt)));
  bool face;
  t.foot = closest_on_triangle(c, A, B, C, face);
  const v3 e = c - t.foot;
  const float side = dot(e, nt), d = sqrtf(dot(e, e));
  const bool edge_ok = side > 0.f && d > 1e-6f;
  // within r/4 of an edge or vertex the direction to it is dominated by rounding: the
  // contact then takes the face normal
  const bool flat = face || d <= 0.25f * r;
  t.nrm = flat ? nt : e * (1.f / fmaxf(d, 1e-6f));
  t.dist = (flat ? side : d) - r;
  t.usable = face || edge_ok;
  t.face = face;
  return t;
}

// sphere column a against grid cell q (0..3) of the 2 x 2 block around the grid vertex
// nearest its centre, the cell split into two triangles along its (i, j)-(i+1, j+1)
// diagonal: one of the two triangles' usable contacts
__device__ Cand eval_heightfield_cell(const ColCtx& C, const float* __restrict__ hgt, int a, int q) {
  Cand o;
  const ofb_collider_desc& d = C.d;
  o.s0 = C.flat_id[C.col(a)];
  o.s1 = C.flat_id[C.col(d.hf_col)];
  const tf Ba = C.body_pose(a);
  const v3 c = C.shape_pose(a).p;
  const float r = C.sz(a).x;
  const int nx = d.hf_nx, ny = d.hf_ny;
  const float sx = d.hf_spacing[0], sy = d.hf_spacing[1], ox = d.hf_origin[0], oy = d.hf_origin[1];
  const int I = (int)floorf((c.x - ox) / sx + 0.5f), J = (int)floorf((c.y - oy) / sy + 0.5f);
  const int ci = I - 1 + (q & 1), cj = J - 1 + (q >> 1);
  const bool inside = ci >= 0 && ci <= nx - 2 && cj >= 0 && cj <= ny - 2;
  const int i0 = min(max(ci, 0), nx - 2), j0 = min(max(cj, 0), ny - 2);
  auto vert = [&](int i, int j) { return mk(ox + (float)i * sx, oy + (float)j * sy, hgt[j * nx + i]); };
  const v3 P00 = vert(i0, j0), P10 = vert(i0 + 1, j0), P01 = vert(i0, j0 + 1), P11 = vert(i0 + 1, j0 + 1);
  const TriContact ta = tri_contact(c, r, P00, P10, P11), tb = tri_contact(c, r, P00, P11, P01);
  // a face contact wins over an edge or vertex one; between two of a kind the nearer,
  // near-ties going to the first triangle so that rounding does not pick between two
  // distinct feet
  const bool use_a =
      ta.usable && (!tb.usable || (ta.face && !tb.face) || (ta.face == tb.face && ta.dist <= tb.dist + 1e-6f));
  const TriContact& t = use_a ? ta : tb;
  const v3 pw = c - t.nrm * r;
  o.valid = inside && (ta.usable || tb.usable) && t.dist < d.gap;
  o.p0 = qrot_inv(Ba.q, pw - Ba.p);
  o.p1 = t.foot;
  o.n = t.nrm * -1.f;
  return o;
}

// sphere column a against the heightfield (bilinear surface below the centre)
__device__ Cand eval_heightfield(const ColCtx& C, const float* __restrict__ hgt, int a) {
  Cand o;
  const ofb_collider_desc& d = C.d;
  o.s0 = C.flat_id[C.col(a)];
  o.s1 = C.flat_id[C.col(d.hf_col)];
  const tf Ba = C.body_pose(a);
  const v3 c = C.shape_pose(a).p;
  const float r = C.sz(a).x;
  const int nx = d.hf_nx, ny = d.hf_ny;
  const float u = fminf(fmaxf((c.x - d.hf_origin[0]) / d.hf_spacing[0], 0.f), (float)(nx - 1));
  const float v = fminf(fmaxf((c.y - d.hf_origin[1]) / d.hf_spacing[1], 0.f), (float)(ny - 1));
  const int i = min((int)floorf(u), nx - 2), j = min((int)floorf(v), ny - 2);
  const float fu = u - (float)i, fv = v - (float)j;
  const float h00 = hgt[j * nx + i], h10 = hgt[j * nx + i + 1], h01 = hgt[(j + 1) * nx + i],
              h11 = hgt[(j + 1) * nx + i + 1];
  const float z = h00 * (1.f - fu) * (1.f - fv) + h10 * fu * (1.f - fv) + h01 * (1.f - fu) * fv + h11 * fu * fv;
  const float gx = ((h10 - h00) * (1.f - fv) + (h11 - h01) * fv) / d.hf_spacing[0];
  const float gy = ((h01 - h00) * (1.f - fu) + (h11 - h10) * fu) / d.hf_spacing[1];
  v3 nrm = mk(-gx, -gy, 1.f);
  nrm = nrm * (1.f / sqrtf(dot(nrm, nrm)));
  const float dist = dot(c - mk(c.x, c.y, z), nrm) - r;
  const v3 pw = c - nrm * r;
  o.valid = dist < d.gap;
  o.p0 = qrot_inv(Ba.q, pw - Ba.p);
  o.p1 = pw - nrm * dist;
  o.n = nrm * -1.f;
  return o;
}

// Can any candidate of pair (a, b) be a contact?  Every sample point of shape a lies
// within bound_radius of its centre (for a sphere the candidate is the centre itself, its
// radius enters the thresholds), and a contact needs its signed distance below the gap
// (plane) or every box-frame excess below max(gap (+ radius), slack) (box).  `eps` covers
// the rounding of the sample points, so the test never drops a contact.
__device__ bool pair_may_touch(const ColCtx& C, int a, int b) {
  const int ga = C.geo[C.col(a)], gb = C.geo[C.col(b)];
  const tf Xa = C.shape_pose(a), Xb = C.shape_pose(b);
  const v3 h = C.sz(a);
  float R = 0.f, r = 0.f;
  if (ga == CGEO_BOX) R = sqrtf(dot(h, h));
  else if (ga == CGEO_CYLINDER) R = sqrtf(h.x * h.x + h.y * h.y);
  else r = h.x;  // sphere
  const float eps = 1e-3f + 1e-5f * (sqrtf(dot(Xa.p, Xa.p)) + sqrtf(dot(Xb.p, Xb.p)));
  const v3 d = Xa.p - Xb.p;
  if (gb == CGEO_PLANE) {
    const v3 nrm = qrot(Xb.q, mk(0.f, 0.f, 1.f));
    return dot(d, nrm) - R - r < C.d.gap + eps;
  }
  const v3 q = qrot_inv(Xb.q, d), hb = C.sz(b);
  const float m = fmaxf(C.d.gap, C.d.lateral_slack) + r;
  const v3 e = mk(fmaxf(fabsf(q.x) - hb.x - m, 0.f), fmaxf(fabsf(q.y) - hb.y - m, 0.f),
                  fmaxf(fabsf(q.z) - hb.z - m, 0.f));
  return sqrtf(dot(e, e)) < R + eps;
}

// block-wide exclusive prefix of v in thread order and the block total
__device__ __forceinline__ int block_scan(int v, int* wave_tot, int& total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) wave_tot[wave] = x;
  __syncthreads();
  int before = 0;
  total = 0;
  for (int k = 0; k < kColThreads / 64; ++k) {
    if (k < wave) before += wave_tot[k];
    total += wave_tot[k];
  }
  __syncthreads();
  return before + x - v;
}

struct ColArgs {
  ofb_collider_desc d;
  const int32_t *geo, *body, *flat_id, *pairs, *pair_off, *hf_spheres;
  const float *xform, *size, *hgt;
};

constexpr int kMaxCylSamples = 256;
constexpr int kMaxColPairs = 8192;  // candidate-range table in LDS: (pairs + 1) ints

// Per-world set-up shared by both passes: the rim angle table and the compacted candidate
// ranges of the pairs that may touch (poff[p] .. poff[p+1]); returns their total.
__device__ int world_setup(const ColArgs& A, const ColCtx& C, int* poff, float* ang_cs, float* ang_sn,
                           int* wave_tot) {
  const int P = A.d.num_pairs, k = A.d.cylinder_samples;
  for (int j = threadIdx.x; j < k; j += kColThreads) {
    const float ang = (float)j * (float)(6.283185307179586 / (double)k);
    ang_cs[j] = cosf(ang);
    ang_sn[j] = sinf(ang);
  }
  int run = 0;
  for (int p0 = 0; p0 < P; p0 += kColThreads) {
    const int p = p0 + threadIdx.x;
    int n = 0;
    if (p < P) {
      n = A.pair_off[p + 1] - A.pair_off[p];
      if (!pair_may_touch(C, A.pairs[2 * p], A.pairs[2 * p + 1])) n = 0;
    }
    int tot;
    const int ex = block_scan(n, wave_tot, tot);
    if (p < P) poff[p] = run + ex;
    run += tot;
  }
  if (threadIdx.x == 0) poff[P] = run;
  __syncthreads();
  return run;
}

__device__ __forceinline__ Cand eval_candidate(const ColArgs& A, const ColCtx& C, const int* poff, int active,
                                               const float* ang_cs, const float* ang_sn, int c) {
  if (c >= active) {
    if (A.d.hf_points == 4) return eval_heightfield_cell(C, A.hgt, A.hf_spheres[(c - active) >> 2], (c - active) & 3);
    return eval_heightfield(C, A.hgt, A.hf_spheres[c - active]);
  }
  int lo = 0, hi = A.d.num_pairs;  // largest p with poff[p] <= c (then c < poff[p + 1])
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (poff[mid] <= c) lo = mid;
    else hi = mid;
  }
  return eval_pair_point(C, A.pairs[2 * lo], A.pairs[2 * lo + 1], c - poff[lo], ang_cs, ang_sn);
}

__global__ void __launch_bounds__(kColThreads) collide_count_kernel(ColArgs A, const float* body_q,
                                                                    int32_t* world_count) {
  extern __shared__ int poff[];
  __shared__ int wave_tot[kColThreads / 64];
  __shared__ float ang_cs[kMaxCylSamples], ang_sn[kMaxCylSamples];
  const ColCtx C{A.d, A.geo, A.body, A.flat_id, A.xform, A.size, body_q, (int)blockIdx.x};
  const int active = world_setup(A, C, poff, ang_cs, ang_sn, wave_tot);
  const int cand = active + A.d.num_hf_spheres * A.d.hf_points;
  int count = 0;
  for (int c0 = 0; c0 < cand; c0 += kColThreads) {
    const int c = c0 + threadIdx.x;
    const bool v = c < cand && eval_candidate(A, C, poff, active, ang_cs, ang_sn, c).valid;
    int tot;
    block_scan(v ? 1 : 0, wave_tot, tot);
    count += tot;
  }
  if (threadIdx.x == 0) world_count[blockIdx.x] = count;
}

// exclusive scan of the world counts (one workgroup), total capped at `cap` into count[0]
__global__ void __launch_bounds__(1024) collide_scan_kernel(int W, const int32_t* world_count, int32_t* offset,
                                                            int cap, int32_t* count_out) {
  __shared__ int part[1024];
  const int t = threadIdx.x, per = (W + 1023) / 1024;
  int s = 0;
  for (int i = t * per; i < min(W, (t + 1) * per); ++i) s += world_count[i];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = t > 0 ? part[t - 1] : 0;
  for (int i = t * per; i < min(W, (t + 1) * per); ++i) {
    offset[i] = run;
    run += world_count[i];
  }
  if (t == 1023) {
    offset[W] = part[1023];
    count_out[0] = min(part[1023], cap);
  }
}

__global__ void __launch_bounds__(kColThreads) collide_write_kernel(ColArgs A, const float* body_q,
                                                                    const int32_t* offset, int cap,
                                                                    ofb_contact_buffers out) {
  extern __shared__ int poff[];
  __shared__ int wave_tot[kColThreads / 64];
  __shared__ float ang_cs[kMaxCylSamples], ang_sn[kMaxCylSamples];
  const ColCtx C{A.d, A.geo, A.body, A.flat_id, A.xform, A.size, body_q, (int)blockIdx.x};
  const int active = world_setup(A, C, poff, ang_cs, ang_sn, wave_tot);
  const int cand = active + A.d.num_hf_spheres * A.d.hf_points;
  int base = offset[blockIdx.x];
  for (int c0 = 0; c0 < cand; c0 += kColThreads) {
    const int c = c0 + threadIdx.x;
    Cand k;
    k.valid = false;
    if (c < cand) k = eval_candidate(A, C, poff, active, ang_cs, ang_sn, c);
    int tot;
    const int r = block_scan(k.valid ? 1 : 0, wave_tot, tot);
    const int pos = base + r;
    if (k.valid && pos < cap) {
      out.point0[pos * 3 + 0] = k.p0.x; out.point0[pos * 3 + 1] = k.p0.y; out.point0[pos * 3 + 2] = k.p0.z;
      out.point1[pos * 3 + 0] = k.p1.x; out.point1[pos * 3 + 1] = k.p1.y; out.point1[pos * 3 + 2] = k.p1.z;
      out.normal[pos * 3 + 0] = k.n.x; out.normal[pos * 3 + 1] = k.n.y; out.normal[pos * 3 + 2] = k.n.z;
      out.shape0[pos] = k.s0;
      out.shape1[pos] = k.s1;
      out.margin0[pos] = 0.f;
      out.margin1[pos] = 0.f;
    }
    base += tot;
  }
}

template <class T>
int up(T** dst, const T* src, size_t n) {
  *dst = nullptr;
  if (n == 0) return 0;
  if (hipMalloc((void**)dst, n * sizeof(T)) != hipSuccess) return -4;
  if (src && hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return -4;
  return 0;
}

}  // namespace

extern "C" {

int ofb_collider_destroy(ofb_collider* c) {
  if (!c) return 0;
  void* ptrs[] = {c->pairs, c->pair_off, c->geo, c->body, c->flat_id, c->hf_spheres, c->xform, c->size,
                  c->hf_heights, c->world_count, c->world_offset};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete c;
  return 0;
}

int ofb_collider_create(const ofb_collider_desc* d, ofb_collider** out) {
  if (!d || !out) { ofb::set_error("null argument"); return -1; }
  *out = nullptr;
  const int W = d->num_worlds, SC = d->shape_cols, P = d->num_pairs;
  if (W < 1 || d->bodies_per_world < 1 || SC < 1 || P < 0 || d->max_contacts_per_world < 1 ||
      d->cylinder_samples < 1 || (P > 0 && !d->pairs) || !d->shape_geo || !d->shape_body || !d->shape_xform ||
      !d->shape_size || !d->flat_shape_id) {
    ofb::set_error("invalid collider description");
    return -2;
  }
  if (d->cylinder_samples > kMaxCylSamples || P > kMaxColPairs) {
    ofb::set_error("collider limits: cylinder_samples <= %d, num_pairs <= %d", kMaxCylSamples, kMaxColPairs);
    return -2;
  }
  if (d->num_hf_spheres > 0 && d->hf_points != 1 && d->hf_points != 4) {
    ofb::set_error("hf_points must be 1 or 4");
    return -2;
  }
  if (d->num_hf_spheres > 0 && (d->hf_col < 0 || d->hf_col >= SC || d->hf_nx < 2 || d->hf_ny < 2 ||
                                !d->hf_heights || !d->hf_spheres)) {
    ofb::set_error("invalid heightfield description");
    return -2;
  }
  // candidates per pair (collide.py: a cylinder or sphere against a plane is one point)
  std::vector<int32_t> off(P + 1, 0);
  for (int p = 0; p < P; ++p) {
    const int a = d->pairs[2 * p], b = d->pairs[2 * p + 1];
    if (a < 0 || a >= SC || b < 0 || b >= SC) { ofb::set_error("pair %d outside the shape columns", p); return -2; }
    const int ga = d->shape_geo[a], gb = d->shape_geo[b];
    if (gb != CGEO_PLANE && gb != CGEO_BOX) { ofb::set_error("pair %d: shape1 must be a plane or a box", p); return -2; }
    int n = 1;
    if (ga == CGEO_BOX) n = 8;
    else if (ga == CGEO_CYLINDER) n = gb == CGEO_PLANE ? 1 : 3 * d->cylinder_samples;
    else if (ga != CGEO_SPHERE) { ofb::set_error("pair %d: unsupported shape0 geometry %d", p, ga); return -2; }
    off[p + 1] = off[p] + n;
  }
  ofb_collider* c = new ofb_collider();
  c->d = *d;
  // the kernels read the device copies below; drop the caller's host pointers
  c->d.pairs = c->d.shape_geo = c->d.shape_body = c->d.flat_shape_id = c->d.hf_spheres = nullptr;
  c->d.shape_xform = c->d.shape_size = c->d.hf_heights = nullptr;
  const size_t WS = (size_t)W * SC;
  int rc = 0;
  rc |= up(&c->pairs, d->pairs, (size_t)2 * P);
  rc |= up(&c->pair_off, off.data(), off.size());
  rc |= up(&c->geo, d->shape_geo, WS);
  rc |= up(&c->body, d->shape_body, WS);
  rc |= up(&c->flat_id, d->flat_shape_id, WS);
  rc |= up(&c->xform, d->shape_xform, WS * 7);
  rc |= up(&c->size, d->shape_size, WS * 3);
  if (d->num_hf_spheres > 0) {
    rc |= up(&c->hf_spheres, d->hf_spheres, (size_t)d->num_hf_spheres);
    rc |= up(&c->hf_heights, d->hf_heights, (size_t)d->hf_nx * d->hf_ny);
  }
  rc |= up<int32_t>(&c->world_count, nullptr, (size_t)W);
  rc |= up<int32_t>(&c->world_offset, nullptr, (size_t)W + 1);
  if (rc) {
    ofb_collider_destroy(c);
    ofb::set_error("hipMalloc failed (collider)");
    return -4;
  }
  *out = c;
  return 0;
}

int ofb_collide(ofb_collider* c, const float* body_q, const ofb_contact_buffers* out, void* stream) {
  if (!c || !body_q || !out || !out->count || !out->point0 || !out->point1 || !out->normal || !out->shape0 ||
      !out->shape1 || !out->margin0 || !out->margin1) {
    ofb::set_error("null argument");
    return -1;
  }
  const int W = c->d.num_worlds;
  const int cap = W * c->d.max_contacts_per_world;
  if (out->max_contacts < cap) {
    ofb::set_error("contact buffers hold %d slots, the collider emits up to %d", out->max_contacts, cap);
    return -2;
  }
  hipStream_t st = (hipStream_t)stream;
  const ColArgs A{c->d, c->geo, c->body, c->flat_id, c->pairs, c->pair_off, c->hf_spheres, c->xform, c->size,
                  c->hf_heights};
  const size_t lds = (size_t)(c->d.num_pairs + 1) * sizeof(int);
  hipLaunchKernelGGL(collide_count_kernel, dim3(W), dim3(kColThreads), lds, st, A, body_q, c->world_count);
  hipLaunchKernelGGL(collide_scan_kernel, dim3(1), dim3(1024), 0, st, W, c->world_count, c->world_offset, cap,
                     out->count);
  hipLaunchKernelGGL(collide_write_kernel, dim3(W), dim3(kColThreads), lds, st, A, body_q, c->world_offset, cap,
                     *out);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ofb::set_error("collider launch failed: %s", hipGetErrorString(e));
    return -4;
  }
  return 0;
}

}  // extern "C"
