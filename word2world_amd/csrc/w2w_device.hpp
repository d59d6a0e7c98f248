// Device-side math shared by the rasterizer kernels (gfx950, wave64).
//
// Compiled with -ffp-contract=off: every multiply and add rounds on its own,
// and fused multiply-adds are explicit __builtin_fmaf.  The operation order
// of the geometry, radius and blend arithmetic is fixed so that tile counts,
// keys, ranges, n_contrib and the final transmittance are reproducible bit
// for bit on any IEEE float32 implementation that follows the same order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define W2W_DEV __device__ __forceinline__
#define W2W_HD __host__ __device__ __forceinline__

namespace w2w {

// exp(x) for x in [-87, 0]: Cody-Waite split x = k ln2 + r, degree-6 Taylor
// polynomial for e^r, scaled by 2^k.  ~1 ulp; deterministic bits.
W2W_DEV float exp_neg(float x) {
  float kf = __builtin_rintf(x * 0x1.715476p+0f);
  float r = __builtin_fmaf(kf, -0x1.62e4p-1f, x);
  r = __builtin_fmaf(kf, -0x1.7f7d1cp-20f, r);
  float p = 0x1.6c16c2p-10f;
  p = __builtin_fmaf(p, r, 0x1.111112p-7f);
  p = __builtin_fmaf(p, r, 0x1.555556p-5f);
  p = __builtin_fmaf(p, r, 0x1.555556p-3f);
  p = __builtin_fmaf(p, r, 0x1p-1f);
  p = __builtin_fmaf(p, r, 1.0f);
  p = __builtin_fmaf(p, r, 1.0f);
  return __builtin_ldexpf(p, (int)kf);
}

// ln(x) for finite normal x > 0: x = m 2^e, m in [sqrt(1/2), sqrt(2)),
// ln m = 2 atanh((m-1)/(m+1)) by its odd series to s^11.
W2W_DEV float log_pos(float x) {
  uint32_t ix = __float_as_uint(x);
  int e = (int)((ix >> 23) & 0xffu) - 127;
  float m = __uint_as_float((ix & 0x7fffffu) | 0x3f800000u);
  if (m > 0x1.6a09e6p+0f) {
    m = m * 0.5f;
    e = e + 1;
  }
  float f = m - 1.0f;
  float s = f / (2.0f + f);
  float z = s * s;
  float p = 0x1.745d18p-4f;
  p = __builtin_fmaf(p, z, 0x1.c71c72p-4f);
  p = __builtin_fmaf(p, z, 0x1.24924ap-3f);
  p = __builtin_fmaf(p, z, 0x1.99999ap-3f);
  p = __builtin_fmaf(p, z, 0x1.555556p-2f);
  p = __builtin_fmaf(p, z, 1.0f);
  float lnm = (2.0f * s) * p;
  float ef = (float)e;
  return __builtin_fmaf(ef, 0x1.62e4p-1f, __builtin_fmaf(ef, 0x1.7f7d1cp-20f, lnm));
}

// Gaussian exponent -0.5 (a dx^2 + c dy^2) - b dx dy in a fixed order.
W2W_HD float power_of(float ca, float cb, float cc, float dx, float dy) {
  float q = __builtin_fmaf(ca * dx, dx, (cc * dy) * dy);
  return __builtin_fmaf(-0.5f, q, -((cb * dx) * dy));
}

// Reference SH constants (gaussianForwardPass.cu:11-39; coefficient 6 = 0.3).
constexpr float SH_C0 = 0.28209479177387814f;
constexpr float SH_C1 = 0.4886025119029199f;
constexpr float SH_C2A = 1.0925484305920792f;
constexpr float SH_C2E = 0.5462742152960396f;
constexpr float SH_C26 = 0.3f;
constexpr float SH_K9 = -0.5900435899266435f;
constexpr float SH_K10 = 2.890611442640554f;
constexpr float SH_K11 = -0.4570457994644658f;
constexpr float SH_K12 = 0.3731763325901154f;
constexpr float SH_K13 = -0.4570457994644658f;
constexpr float SH_K14 = 1.445305721320277f;
constexpr float SH_K15 = -0.5900435899266435f;

W2W_DEV void sh_basis(int D, float x, float y, float z, float* Y) {
  Y[0] = SH_C0;
  if (D < 1) return;
  Y[1] = -SH_C1 * y;
  Y[2] = SH_C1 * z;
  Y[3] = -SH_C1 * x;
  if (D < 2) return;
  float xx = x * x, yy = y * y, zz = z * z;
  float xy = x * y, yz = y * z, xz = x * z;
  Y[4] = SH_C2A * xy;
  Y[5] = -SH_C2A * yz;
  Y[6] = SH_C26 * ((2.0f * zz - xx) - yy);
  Y[7] = -SH_C2A * xz;
  Y[8] = SH_C2E * (xx - yy);
  if (D < 3) return;
  Y[9] = (SH_K9 * y) * (3.0f * xx - yy);
  Y[10] = ((SH_K10 * x) * y) * z;
  Y[11] = (SH_K11 * y) * ((4.0f * zz - xx) - yy);
  Y[12] = (SH_K12 * z) * ((2.0f * zz - 3.0f * xx) - 3.0f * yy);
  Y[13] = (SH_K13 * x) * ((4.0f * zz - xx) - yy);
  Y[14] = (SH_K14 * z) * (xx - yy);
  Y[15] = (SH_K15 * x) * (xx - 3.0f * yy);
}

W2W_DEV void sh_basis_grad(int D, float x, float y, float z, float* gx, float* gy, float* gz) {
#pragma unroll
  for (int k = 0; k < 16; k++) gx[k] = gy[k] = gz[k] = 0.0f;
  if (D < 1) return;
  gy[1] = -SH_C1;
  gz[2] = SH_C1;
  gx[3] = -SH_C1;
  if (D < 2) return;
  float xx = x * x, yy = y * y, zz = z * z;
  gx[4] = SH_C2A * y; gy[4] = SH_C2A * x;
  gy[5] = -SH_C2A * z; gz[5] = -SH_C2A * y;
  gx[6] = SH_C26 * (-2.0f * x); gy[6] = SH_C26 * (-2.0f * y); gz[6] = SH_C26 * (4.0f * z);
  gx[7] = -SH_C2A * z; gz[7] = -SH_C2A * x;
  gx[8] = SH_C2E * (2.0f * x); gy[8] = SH_C2E * (-2.0f * y);
  if (D < 3) return;
  gx[9] = SH_K9 * (6.0f * x * y); gy[9] = SH_K9 * (3.0f * xx - 3.0f * yy);
  gx[10] = SH_K10 * (y * z); gy[10] = SH_K10 * (x * z); gz[10] = SH_K10 * (x * y);
  gx[11] = SH_K11 * (-2.0f * x * y); gy[11] = SH_K11 * ((4.0f * zz - xx) - 3.0f * yy); gz[11] = SH_K11 * (8.0f * y * z);
  gx[12] = SH_K12 * (-6.0f * x * z); gy[12] = SH_K12 * (-6.0f * y * z); gz[12] = SH_K12 * ((6.0f * zz - 3.0f * xx) - 3.0f * yy);
  gx[13] = SH_K13 * ((4.0f * zz - 3.0f * xx) - yy); gy[13] = SH_K13 * (-2.0f * x * y); gz[13] = SH_K13 * (8.0f * x * z);
  gx[14] = SH_K14 * (2.0f * x * z); gy[14] = SH_K14 * (-2.0f * y * z); gz[14] = SH_K14 * (xx - yy);
  gx[15] = SH_K15 * (3.0f * xx - 3.0f * yy); gy[15] = SH_K15 * (-6.0f * x * y);
}

// ---- batch description (device memory) ----
// A batch is a set of scenes (packed Gaussian arrays) and frames (views).
// Work items are (frame, Gaussian) pairs; scene s owns items
// [item_off, item_off + n * n_views) with item = item_off + g * n_views + v
// for Gaussian g and view v (frame frame_off + v), so a Gaussian's views are
// adjacent and, within one frame, items are in Gaussian order.  Tiles of all
// frames share one global index space ([tile_off, tile_off + tiles) per
// frame) and per-pixel outputs one pixel space ([pix_off, pix_off + W*H)).
struct BatchHeader {
  uint32_t n_scenes, n_frames, n_items, n_gaussians, n_tiles, pad0, pad1, pad2;
  uint64_t n_pixels, pad3;
};
struct SceneDesc {
  uint32_t g_off;      // first Gaussian of the scene in the batch's Gaussian order
  uint32_t n;          // Gaussians
  uint32_t item_off;   // first item
  uint32_t frame_off;  // first frame
  uint32_t n_views;    // frames of this scene
  int32_t M, D;        // SH coefficients per Gaussian, active degree
  uint32_t g_base;     // row of the scene's first Gaussian in the parameter and
                       // per-Gaussian gradient arrays (== g_off when packed)
  uint64_t sh_off;     // first float of the scene's SH block [n, M, 3]
  uint64_t pad2;
};
struct FrameDesc {
  float V[16];         // world->camera, column-major (Renderer.py:135)
  float F[16];         // full projection P*W2C, column-major
  float cp[3], fx;     // caller's campos, focal x
  float bg[3], fy;     // background, focal y
  int32_t W, H, tiles_x, tiles_y;
  uint32_t tile_off, scene, view, flags;  // FRAME_* bits
  uint64_t pix_off;    // first pixel in the packed per-pixel outputs
  uint64_t target_off; // first float of this frame's [3,H,W] loss target
};
constexpr uint32_t FRAME_BACKWARD = 1u;  // gradients wanted for this frame
// Emit a Gaussian only into the tiles of its rectangle that its cull extent
// box meets (the rectangle of the reference rule intersected with that box):
// fewer instances, identical images, final T and gradients; K, the keys, the
// ranges and n_contrib (an index into the tile's list) then differ from the
// reference's.
constexpr uint32_t FRAME_TIGHT = 2u;
constexpr size_t TABLE_HEADER_BYTES = 64;

W2W_DEV const SceneDesc* table_scenes(const void* table) {
  return reinterpret_cast<const SceneDesc*>(static_cast<const char*>(table) + TABLE_HEADER_BYTES);
}
W2W_DEV const FrameDesc* table_frames(const void* table, uint32_t scenes_cap) {
  return reinterpret_cast<const FrameDesc*>(static_cast<const char*>(table) + TABLE_HEADER_BYTES +
                                            (size_t)scenes_cap * sizeof(SceneDesc));
}

// largest index k in [0, n) with key(k) <= x (keys non-decreasing, key(0) <= x)
template <typename F>
W2W_DEV uint32_t upper_index(uint32_t n, uint32_t x, F key) {
  uint32_t lo = 0, hi = n - 1;
  while (lo < hi) {
    uint32_t mid = (lo + hi + 1) >> 1;
    if (key(mid) <= x) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Camera constants of one frame, read once per thread (uniform addresses).
struct Cam {
  float V[16];
  float F[16];
  float cp[3];
  float fx, fy, limx, limy;
  int W, H;
};

W2W_DEV void load_cam(Cam& c, const FrameDesc& fd) {
#pragma unroll
  for (int k = 0; k < 16; k++) {
    c.V[k] = fd.V[k];
    c.F[k] = fd.F[k];
  }
  c.cp[0] = fd.cp[0];
  c.cp[1] = fd.cp[1];
  c.cp[2] = fd.cp[2];
  c.fx = fd.fx;
  c.fy = fd.fy;
  c.W = fd.W;
  c.H = fd.H;
  c.limx = (float)fd.W / (2.0f * fd.fx);
  c.limy = (float)fd.H / (2.0f * fd.fy);
}

// Per-Gaussian projection state: the EWA projection of projection.cu:7-154
// with PD-1/PD-2/PD-12/PD-13 applied.
struct Geom {
  bool visible;
  bool clamped;
  float tw, vx, vy, vz;
  float rx, ry, rxc, ryc, s;
  float txc, tyc;
  float jx[3], jy[3];
  float R[9], M[9], Sig[6];
  float cxx, cxy, cyy, det;
  float ca, cb, cc;
  float cx, cy, cz, invW;
  float ndcz;
  float px, py;
};

// R (row-major) of the unit quaternion q = (w, x, y, z) and M = R diag(s),
// the factors of Sigma = M M^T (the same expressions geometry() uses).
W2W_DEV void rot_scale(const float q[4], const float s[3], float R[9], float M[9]) {
  float w = q[0], x = q[1], y = q[2], z = q[3];
  float xx = x * x, yy = y * y, zz = z * z;
  float xy = x * y, xz = x * z, yz = y * z;
  float wx = w * x, wy = w * y, wz = w * z;
  R[0] = 1.0f - 2.0f * (yy + zz); R[1] = 2.0f * (xy - wz);        R[2] = 2.0f * (xz + wy);
  R[3] = 2.0f * (xy + wz);        R[4] = 1.0f - 2.0f * (xx + zz); R[5] = 2.0f * (yz - wx);
  R[6] = 2.0f * (xz - wy);        R[7] = 2.0f * (yz + wx);        R[8] = 1.0f - 2.0f * (xx + yy);
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) M[3 * i + j] = R[3 * i + j] * s[j];
}

// Calls f(k, Y_k, dY_k/dx, dY_k/dy, dY_k/dz) for the real SH basis functions
// of degree <= D at the unit direction (x, y, z), in coefficient order, with
// the expressions of sh_basis / sh_basis_grad; everything is straight-line
// code, so a caller's per-coefficient accumulators stay in registers.
template <int D, typename F>
W2W_DEV void sh_visit(float x, float y, float z, F&& f) {
  f(0, SH_C0, 0.0f, 0.0f, 0.0f);
  if constexpr (D >= 1) {
    f(1, -SH_C1 * y, 0.0f, -SH_C1, 0.0f);
    f(2, SH_C1 * z, 0.0f, 0.0f, SH_C1);
    f(3, -SH_C1 * x, -SH_C1, 0.0f, 0.0f);
  }
  if constexpr (D >= 2) {
    const float xx = x * x, yy = y * y, zz = z * z;
    const float xy = x * y, yz = y * z, xz = x * z;
    f(4, SH_C2A * xy, SH_C2A * y, SH_C2A * x, 0.0f);
    f(5, -SH_C2A * yz, 0.0f, -SH_C2A * z, -SH_C2A * y);
    f(6, SH_C26 * ((2.0f * zz - xx) - yy), SH_C26 * (-2.0f * x), SH_C26 * (-2.0f * y), SH_C26 * (4.0f * z));
    f(7, -SH_C2A * xz, -SH_C2A * z, 0.0f, -SH_C2A * x);
    f(8, SH_C2E * (xx - yy), SH_C2E * (2.0f * x), SH_C2E * (-2.0f * y), 0.0f);
    if constexpr (D >= 3) {
      f(9, (SH_K9 * y) * (3.0f * xx - yy), SH_K9 * (6.0f * x * y), SH_K9 * (3.0f * xx - 3.0f * yy), 0.0f);
      f(10, ((SH_K10 * x) * y) * z, SH_K10 * (y * z), SH_K10 * (x * z), SH_K10 * (x * y));
      f(11, (SH_K11 * y) * ((4.0f * zz - xx) - yy), SH_K11 * (-2.0f * x * y),
        SH_K11 * ((4.0f * zz - xx) - 3.0f * yy), SH_K11 * (8.0f * y * z));
      f(12, (SH_K12 * z) * ((2.0f * zz - 3.0f * xx) - 3.0f * yy), SH_K12 * (-6.0f * x * z),
        SH_K12 * (-6.0f * y * z), SH_K12 * ((6.0f * zz - 3.0f * xx) - 3.0f * yy));
      f(13, (SH_K13 * x) * ((4.0f * zz - xx) - yy), SH_K13 * ((4.0f * zz - 3.0f * xx) - yy),
        SH_K13 * (-2.0f * x * y), SH_K13 * (8.0f * x * z));
      f(14, (SH_K14 * z) * (xx - yy), SH_K14 * (2.0f * x * z), SH_K14 * (-2.0f * y * z), SH_K14 * (xx - yy));
      f(15, (SH_K15 * x) * (xx - 3.0f * yy), SH_K15 * (3.0f * xx - 3.0f * yy), SH_K15 * (-6.0f * x * y), 0.0f);
    }
  }
}

// R, M = R diag(s) and Sigma = M M^T (xx, xy, xz, yy, yz, zz) of a Gaussian.
W2W_DEV void sigma_of(const float q[4], const float s[3], float R[9], float Mm[9], float sig[6]) {
  float w = q[0], x = q[1], y = q[2], z = q[3];
  float xx = x * x, yy = y * y, zz = z * z;
  float xy = x * y, xz = x * z, yz = y * z;
  float wx = w * x, wy = w * y, wz = w * z;
  R[0] = 1.0f - 2.0f * (yy + zz); R[1] = 2.0f * (xy - wz);        R[2] = 2.0f * (xz + wy);
  R[3] = 2.0f * (xy + wz);        R[4] = 1.0f - 2.0f * (xx + zz); R[5] = 2.0f * (yz - wx);
  R[6] = 2.0f * (xz - wy);        R[7] = 2.0f * (yz + wx);        R[8] = 1.0f - 2.0f * (xx + yy);
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) Mm[3 * i + j] = R[3 * i + j] * s[j];
#define W2W_SIGE(i, j) ((Mm[3 * (i) + 0] * Mm[3 * (j) + 0] + Mm[3 * (i) + 1] * Mm[3 * (j) + 1]) + Mm[3 * (i) + 2] * Mm[3 * (j) + 2])
  sig[0] = W2W_SIGE(0, 0); sig[1] = W2W_SIGE(0, 1); sig[2] = W2W_SIGE(0, 2);
  sig[3] = W2W_SIGE(1, 1); sig[4] = W2W_SIGE(1, 2); sig[5] = W2W_SIGE(2, 2);
#undef W2W_SIGE
}

// geometry() with the view-independent Sigma (6 entries) given: everything
// else of the projection, per view.  g.R and g.M are not set.
W2W_DEV void geometry_sig(const Cam& c, const float p[3], const float sig[6], Geom& g) {
  const float* V = c.V;
  g.visible = false;
  float tx = ((V[0] * p[0] + V[4] * p[1]) + V[8] * p[2]) + V[12];
  float ty = ((V[1] * p[0] + V[5] * p[1]) + V[9] * p[2]) + V[13];
  float tz = ((V[2] * p[0] + V[6] * p[1]) + V[10] * p[2]) + V[14];
  g.tw = ((V[3] * p[0] + V[7] * p[1]) + V[11] * p[2]) + V[15];
  g.vx = tx / g.tw;
  g.vy = ty / g.tw;
  g.vz = tz / g.tw;
  if (!(g.vz > 1.0f)) return;
  g.rx = g.vx / g.vz;
  g.ry = g.vy / g.vz;
  float nx = g.rx / c.limx, ny = g.ry / c.limy;
  float sq = nx * nx + ny * ny;
  g.clamped = sq > 1.0f;
  g.s = 1.0f;
  g.rxc = g.rx;
  g.ryc = g.ry;
  if (g.clamped) {
    g.s = 1.0f / __builtin_sqrtf(sq);
    g.rxc = g.rx * g.s;
    g.ryc = g.ry * g.s;
  }
  g.txc = g.rxc * g.vz;
  g.tyc = g.ryc * g.vz;
  float zi = 1.0f / g.vz;
  float zi2 = zi * zi;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    float right = V[4 * k + 0], up = V[4 * k + 1], fwd = V[4 * k + 2];
    g.jx[k] = c.fx * (right * zi - (g.txc * fwd) * zi2);
    g.jy[k] = c.fy * (up * zi - (g.tyc * fwd) * zi2);
  }
#pragma unroll
  for (int k = 0; k < 6; k++) g.Sig[k] = sig[k];
  const float* S6 = g.Sig;
  float v0 = (S6[0] * g.jx[0] + S6[1] * g.jx[1]) + S6[2] * g.jx[2];
  float v1 = (S6[1] * g.jx[0] + S6[3] * g.jx[1]) + S6[4] * g.jx[2];
  float v2 = (S6[2] * g.jx[0] + S6[4] * g.jx[1]) + S6[5] * g.jx[2];
  float w0 = (S6[0] * g.jy[0] + S6[1] * g.jy[1]) + S6[2] * g.jy[2];
  float w1 = (S6[1] * g.jy[0] + S6[3] * g.jy[1]) + S6[4] * g.jy[2];
  float w2 = (S6[2] * g.jy[0] + S6[4] * g.jy[1]) + S6[5] * g.jy[2];
  g.cxx = (g.jx[0] * v0 + g.jx[1] * v1) + g.jx[2] * v2;
  g.cxy = (g.jx[0] * w0 + g.jx[1] * w1) + g.jx[2] * w2;
  g.cyy = (g.jy[0] * w0 + g.jy[1] * w1) + g.jy[2] * w2;
  g.det = g.cxx * g.cyy - g.cxy * g.cxy;
  if (!(__builtin_fabsf(g.det) >= 1e-8f)) return;
  float dinv = 1.0f / g.det;
  g.ca = g.cyy * dinv;
  g.cb = -g.cxy * dinv;
  g.cc = g.cxx * dinv;
  const float* F = c.F;
  g.cx = ((F[0] * p[0] + F[4] * p[1]) + F[8] * p[2]) + F[12];
  g.cy = ((F[1] * p[0] + F[5] * p[1]) + F[9] * p[2]) + F[13];
  g.cz = ((F[2] * p[0] + F[6] * p[1]) + F[10] * p[2]) + F[14];
  float cw = ((F[3] * p[0] + F[7] * p[1]) + F[11] * p[2]) + F[15];
  g.invW = 1.0f / (cw + 1e-5f);
  float ndcx = g.cx * g.invW;
  float ndcy = g.cy * g.invW;
  g.ndcz = g.cz * g.invW;
  g.px = ((ndcx + 1.0f) * 0.5f) * (float)c.W - 0.5f;
  g.py = ((ndcy + 1.0f) * 0.5f) * (float)c.H - 0.5f;
  g.visible = true;
}

W2W_DEV void geometry(const Cam& c, const float p[3], const float s[3], const float q[4], Geom& g) {
  float sig[6];
  sigma_of(q, s, g.R, g.M, sig);
  geometry_sig(c, p, sig, g);
}

W2W_DEV uint32_t tile_coord(float v, uint32_t n) {
  if (!(v > 0.0f)) return 0u;
  if (v >= (float)n) return n;
  return (uint32_t)v;
}

// The splat fields the blend kernels evaluate (48 bytes, 16-byte aligned):
//   a = (px, py, conic_a, conic_b), b = (conic_c, opacity, depth, skip_below),
//   c = (r, g, b, cull extent: two binary16 half-widths packed in 32 bits).
struct __align__(16) Splat {
  float4 a, b, c;
};

// What the preprocess stores per item and the blends gather per instance,
// 64 bytes on a 64-byte boundary so that one gather is one 64-byte HBM
// segment: the splat, the block test's threshold (cull_threshold, below),
// the item's tile rectangle (x0 | y0 << 16, x1 | y1 << 16) and the slot of
// its first backward record (an item-order scan of the tile counts; the
// item's records are contiguous, one per tile of its rectangle in row-major
// order), which together place an instance's backward record without
// further gathers.
struct __align__(64) SplatRec {
  Splat s;
  float q1;           // block-test threshold
  uint32_t rect0, rect1;
  uint32_t item_off;
};

constexpr uint32_t kExtentInf = 0x7C007C00u;  // (+inf, +inf) in binary16

// Half-widths (hx, hy) of an axis-aligned box around the 2D mean outside of
// which the exponent power_of() computes in float is below `skip` at every
// pixel, so a blend kernel may drop the pair before evaluating it.
//
// With Q(d) = a dx^2 + 2 b dx dy + c dy^2 and q0 = -2 skip, the set Q <= q0
// is an ellipse whose bounding box is |dx| <= sqrt(q0 c / det),
// |dy| <= sqrt(q0 a / det), det = a c - b^2.  The float evaluation of
// power_of (dx, dy rounded from the pixel offset, five rounded operations)
// errs by at most 3 u (a dx^2 + c dy^2) <= 3 u kappa Q with
// kappa = (a + c) / lambda_min and u = 2^-24, so a pixel that passes the
// float test satisfies Q <= q0 (1 + 2u) / (1 - 6 u kappa); q0 is widened by
// (1 + 4u) / (1 - 8 u kappa) and the box by a further 2^-9 relative, which
// also covers the double-precision evaluation here, the rounding of the
// pixel offset and the upward binary16 rounding.  Ill-conditioned conics
// (8 u kappa > 0.5) and non-positive-definite ones get an infinite box.
W2W_HD uint32_t cull_extent(float ca, float cb, float cc, float skip) {
  const double a = ca, b = cb, c = cc;
  if (!(a > 0.0) || !(c > 0.0) || !(skip < 0.0f)) return kExtentInf;
  const double det = a * c - b * b;
  if (!(det > 0.0)) return kExtentInf;
  const double h = 0.5 * (a - c);
  const double lmax = 0.5 * (a + c) + __builtin_sqrt(h * h + b * b);
  const double kappa = (a + c) * lmax / det;
  const double u = 0x1p-24;
  const double den = 1.0 - 8.0 * u * kappa;
  if (!(den > 0.5)) return kExtentInf;
  const double q1 = (-2.0 * (double)skip) * (1.0 + 4.0 * u) / den;
  const double grow = 1.0 + 0x1p-9;
  const double hx = __builtin_sqrt(q1 * c / det) * grow + 1e-3;
  const double hy = __builtin_sqrt(q1 * a / det) * grow + 1e-3;
  // binary16 round-to-nearest errs by at most 2^-11 relative: `grow` keeps
  // the stored value above the true half-width; beyond 60000 use +inf
  const uint16_t bx = hx < 60000.0 ? __builtin_bit_cast(uint16_t, (_Float16)(float)hx) : 0x7C00u;
  const uint16_t by = hy < 60000.0 ? __builtin_bit_cast(uint16_t, (_Float16)(float)hy) : 0x7C00u;
  return (uint32_t)bx | ((uint32_t)by << 16);
}

W2W_HD float2 unpack_extent(float w) {
  const uint32_t bits = __builtin_bit_cast(uint32_t, w);
  const _Float16 lo = __builtin_bit_cast(_Float16, (uint16_t)(bits & 0xFFFFu));
  const _Float16 hi = __builtin_bit_cast(_Float16, (uint16_t)(bits >> 16));
  return make_float2((float)lo, (float)hi);
}

// Whether a splat centred at (mx, my) with extent (hx, hy) may touch pixel
// centres in [x0, x1] x [y0, y1].  NaN positions count as touching.
W2W_HD bool extent_touches(float mx, float my, float2 e, float x0, float x1, float y0, float y1) {
  return !((x0 - mx) > e.x || (mx - x1) > e.x || (y0 - my) > e.y || (my - y1) > e.y);
}

// Exact form of the same question: whether the ellipse Q(m - p) <= q1 meets
// the rectangle [x0, x1] x [y0, y1] (pixel centres lie inside it), with q1 the
// rounding-widened threshold of cull_extent.  If the centre is outside the
// rectangle the minimum of the convex Q over it lies on an edge, and on each
// edge it is a clamped one-dimensional quadratic minimum (in double).  Only a
// minimum above q1 (1 + 1e-6) counts as a miss; every degenerate case, NaN
// included, counts as touching.
//
// Split in two: cull_threshold, the per-splat threshold formed once by the
// preprocess (rounded up to float), and ellipse_touches_q per (splat,
// block), which forms the edge minimisers' slopes -b/c and -b/a by float
// division: they move the edge minimiser by |b/c| 2^-24 |dx| at most, which
// raises the computed minimum by a factor of at most 1 + (b^2 / det) 2^-48 <
// 1 + 1e-7 for every conic with a finite threshold (8 u kappa <= 0.5 bounds
// b^2 / det), inside the 1e-6 widening.
W2W_HD float cull_threshold(float ca, float cb, float cc, float skip) {
  const float inf = __builtin_inff();
  const double a = ca, b = cb, c = cc;
  if (!(a > 0.0) || !(c > 0.0) || !(skip < 0.0f)) return inf;
  const double det = a * c - b * b;
  if (!(det > 0.0)) return inf;
  const double h = 0.5 * (a - c);
  const double lmax = 0.5 * (a + c) + __builtin_sqrt(h * h + b * b);
  const double u = 0x1p-24;
  const double den = 1.0 - 8.0 * u * ((a + c) * lmax / det);
  if (!(den > 0.5)) return inf;
  const double q1 = (-2.0 * (double)skip) * (1.0 + 4.0 * u) / den * (1.0 + 1e-6);
  float q1f = (float)q1;
  if ((double)q1f < q1) q1f = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, q1f) + 1u);  // round up
  return q1f;
}

W2W_HD bool ellipse_touches_q(float mx, float my, float ca, float cb, float cc, float q1, float x0, float x1,
                              float y0, float y1) {
  if (!(q1 < __builtin_inff())) return true;
  const double X = mx, Y = my;
  if (!(X < x0 || X > x1 || Y < y0 || Y > y1)) return true;  // centre inside (or NaN)
  // a finite threshold implies a > 0, c > 0
  const double a = ca, b = cb, c = cc, kc = (double)(-cb / cc), ka = (double)(-cb / ca);
  double qmin = 1e300;
  // vertical edges p.x = E, offset dx = X - E, dy in [Y - y1, Y - y0]
  for (int e = 0; e < 2; e++) {
    const double dx = X - (double)(e ? x1 : x0);
    double dy = kc * dx;
    dy = dy < Y - y1 ? Y - y1 : (dy > Y - y0 ? Y - y0 : dy);
    const double q = (a * dx * dx + 2.0 * b * dx * dy) + c * dy * dy;
    qmin = q < qmin ? q : qmin;
  }
  // horizontal edges p.y = E
  for (int e = 0; e < 2; e++) {
    const double dy = Y - (double)(e ? y1 : y0);
    double dx = ka * dy;
    dx = dx < X - x1 ? X - x1 : (dx > X - x0 ? X - x0 : dx);
    const double q = (a * dx * dx + 2.0 * b * dx * dy) + c * dy * dy;
    qmin = q < qmin ? q : qmin;
  }
  return !(qmin > (double)q1);
}

W2W_HD bool ellipse_touches(float mx, float my, float ca, float cb, float cc, float skip, float x0,
                            float x1, float y0, float y1) {
  return ellipse_touches_q(mx, my, ca, cb, cc, cull_threshold(ca, cb, cc, skip), x0, x1, y0, y1);
}

// ---- wave64 helpers ----
W2W_DEV float wave_sum(float v) {
  // butterfly over 64 lanes
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

W2W_DEV uint32_t wave_incl_scan_u32(uint32_t v, int lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

// agent-scope (device-wide, all XCDs) relaxed atomics for look-back words
W2W_DEV uint32_t ld_acquire_dev(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
W2W_DEV void st_release_dev(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace w2w
