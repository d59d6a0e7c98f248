/*
 * w2w_oracle.c -- CPU restatement of the Word2World tile rasterizer.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker.
 * The product path (word2world_amd/) never links, loads or calls it.
 *
 * Every stage restates the reference algorithm of cudaForwardBackwardRenderer/
 * (file:line citations on each function) with the parity decisions PD-1..PD-14
 * applied (generic W,H and tile size; one tile-rect rule for count and
 * emission; culled Gaussians get radius 0 and zero gradients; conic-form alpha;
 * gradients are the exact derivative of the forward defined here).
 *
 * Float order is fixed: compile with -ffp-contract=off, every fused
 * multiply-add is an explicit fmaf().  The HIP kernels use the same order, so
 * tile counts, keys, ranges, n_contrib and final T agree bit for bit.
 *
 * exp and log are project-written (Cody-Waite reduction plus Taylor / atanh
 * series) so that the GPU and this file compute identical bits; the reference
 * used CUDA expf/logf whose bits are unreachable here.
 *
 * Pinning (tests/test_oracle_*.py): closed forms for Sigma, cov2D, pixel mean,
 * SH at canonical directions, single- and two-Gaussian blends; brute-force tile
 * enumeration; numpy cumsum / lexsort / searchsorted for scan, sort, ranges;
 * fp64 autograd and central finite differences for the backward.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------ */
/* project-written exp / log                                           */
/* ------------------------------------------------------------------ */

/* exp(x) for x in [-87, 0]: Cody-Waite split x = k ln2 + r, |r| <= ln2/2,
 * degree-6 Taylor polynomial for e^r, scaled by 2^k.  ~1 ulp. */
static float w2w_expf(float x) {
  float kf = rintf(x * 0x1.715476p+0f);
  float r = fmaf(kf, -0x1.62e4p-1f, x);
  r = fmaf(kf, -0x1.7f7d1cp-20f, r);
  float p = 0x1.6c16c2p-10f;
  p = fmaf(p, r, 0x1.111112p-7f);
  p = fmaf(p, r, 0x1.555556p-5f);
  p = fmaf(p, r, 0x1.555556p-3f);
  p = fmaf(p, r, 0x1p-1f);
  p = fmaf(p, r, 1.0f);
  p = fmaf(p, r, 1.0f);
  return ldexpf(p, (int)kf);
}

/* ln(x) for finite normal x > 0: x = m 2^e with m in [sqrt(1/2), sqrt(2)),
 * ln m = 2 atanh(s), s = (m-1)/(m+1), odd series to s^11.  ~1 ulp. */
static float w2w_logf(float x) {
  uint32_t ix;
  memcpy(&ix, &x, 4);
  int e = (int)((ix >> 23) & 0xffu) - 127;
  uint32_t mb = (ix & 0x7fffffu) | 0x3f800000u;
  float m;
  memcpy(&m, &mb, 4);
  if (m > 0x1.6a09e6p+0f) {
    m = m * 0.5f;
    e = e + 1;
  }
  float f = m - 1.0f;
  float s = f / (2.0f + f);
  float z = s * s;
  float p = 0x1.745d18p-4f;
  p = fmaf(p, z, 0x1.c71c72p-4f);
  p = fmaf(p, z, 0x1.24924ap-3f);
  p = fmaf(p, z, 0x1.99999ap-3f);
  p = fmaf(p, z, 0x1.555556p-2f);
  p = fmaf(p, z, 1.0f);
  float lnm = (2.0f * s) * p;
  float ef = (float)e;
  return fmaf(ef, 0x1.62e4p-1f, fmaf(ef, 0x1.7f7d1cp-20f, lnm));
}

EXPORT float oracle_expf(float x) { return w2w_expf(x); }
EXPORT float oracle_logf(float x) { return w2w_logf(x); }

/* ------------------------------------------------------------------ */
/* camera                                                              */
/* ------------------------------------------------------------------ */

typedef struct {
  int W, H, ts;      /* image size and tile edge in pixels */
  float fx, fy;      /* focal lengths in pixels */
  float view[16];    /* world->camera, stored as in Renderer.py (column-major) */
  float proj[16];    /* full projection P*W2C, same storage */
  float campos[3];   /* the caller's campos (GaussianTrainer passes -c2w[:3,3]) */
  float bg[3];
} OCam;

static void ocam_fill(OCam* c, const float* view, const float* proj, const float* campos,
                      const float* bg, int W, int H, int ts, float fx, float fy) {
  memcpy(c->view, view, 64);
  memcpy(c->proj, proj, 64);
  memcpy(c->campos, campos, 12);
  memcpy(c->bg, bg, 12);
  c->W = W; c->H = H; c->ts = ts; c->fx = fx; c->fy = fy;
}

/* ------------------------------------------------------------------ */
/* SH basis, reference constants (gaussianForwardPass.cu:11-39, 199)   */
/* Coefficient 6 uses 0.3 (gaussianForwardPass.cu:23, PD-8).           */
/* ------------------------------------------------------------------ */
#define SH_C0 0.28209479177387814f
#define SH_C1 0.4886025119029199f
#define SH_C2A 1.0925484305920792f
#define SH_C2E 0.5462742152960396f
#define SH_C26 0.3f
#define SH_K9 (-0.5900435899266435f)
#define SH_K10 2.890611442640554f
#define SH_K11 (-0.4570457994644658f)
#define SH_K12 0.3731763325901154f
#define SH_K13 (-0.4570457994644658f)
#define SH_K14 1.445305721320277f
#define SH_K15 (-0.5900435899266435f)

static void sh_basis(int D, float x, float y, float z, float* Y) {
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

/* gradient of each basis function w.r.t. the unit direction (x,y,z) */
static void sh_basis_grad(int D, float x, float y, float z, float* gx, float* gy, float* gz) {
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

/* ------------------------------------------------------------------ */
/* per-Gaussian geometry shared by forward and backward                */
/* ------------------------------------------------------------------ */
typedef struct {
  int visible;
  float tx, ty, tz, tw;        /* view-space point before the /w divide */
  float vx, vy, vz;            /* view-space point after /w */
  float rx, ry, rxc, ryc, s;   /* ray-space point, clamped copy, clamp scale */
  int clamped;
  float txc, tyc;              /* clamped view x,y used by the Jacobian */
  float jx[3], jy[3];          /* world-space Jacobian rows */
  float R[9], M[9], Sig[6];    /* rotation, R*S, packed covariance */
  float cxx, cxy, cyy, det;    /* 2D covariance */
  float ca, cb, cc;            /* conic */
  float cx, cy, cz, cw, invW;  /* clip-space point */
  float ndcx, ndcy, ndcz;
  float px, py;                /* pixel-space mean */
} OGeom;

/* projection.cu:7-154 restated (PD-1, PD-2, PD-12, PD-13). */
static void o_geometry(const OCam* c, const float* p, const float* s, const float* q, OGeom* g) {
  const float* V = c->view;
  g->visible = 0;
  /* view transform, glm mat4*vec4 column order (projection.cu:71-74) */
  g->tx = ((V[0] * p[0] + V[4] * p[1]) + V[8] * p[2]) + V[12];
  g->ty = ((V[1] * p[0] + V[5] * p[1]) + V[9] * p[2]) + V[13];
  g->tz = ((V[2] * p[0] + V[6] * p[1]) + V[10] * p[2]) + V[14];
  g->tw = ((V[3] * p[0] + V[7] * p[1]) + V[11] * p[2]) + V[15];
  g->vx = g->tx / g->tw;
  g->vy = g->ty / g->tw;
  g->vz = g->tz / g->tw;
  if (!(g->vz > 1.0f)) return; /* near cull, projection.cu:75-77 (NaN culled too) */

  /* radial FOV clamp (projection.cu:78-92); limit = tan(fov/2) = W/(2 f) */
  float limx = (float)c->W / (2.0f * c->fx);
  float limy = (float)c->H / (2.0f * c->fy);
  g->rx = g->vx / g->vz;
  g->ry = g->vy / g->vz;
  float nx = g->rx / limx, ny = g->ry / limy;
  float sq = nx * nx + ny * ny;
  g->clamped = sq > 1.0f;
  g->s = 1.0f;
  g->rxc = g->rx;
  g->ryc = g->ry;
  if (g->clamped) {
    g->s = 1.0f / sqrtf(sq);
    g->rxc = g->rx * g->s;
    g->ryc = g->ry * g->s;
  }
  g->txc = g->rxc * g->vz;
  g->tyc = g->ryc * g->vz;

  /* Jacobian rows in world space (projection.cu:94-101); rows of W2C */
  float zi = 1.0f / g->vz;
  float zi2 = zi * zi;
  for (int k = 0; k < 3; k++) {
    float right = V[4 * k + 0], up = V[4 * k + 1], fwd = V[4 * k + 2];
    g->jx[k] = c->fx * (right * zi - (g->txc * fwd) * zi2);
    g->jy[k] = c->fy * (up * zi - (g->tyc * fwd) * zi2);
  }

  /* Sigma = (R S)(R S)^T, quaternion (w,x,y,z) (projection.cu:37-69) */
  float w = q[0], x = q[1], y = q[2], z = q[3];
  float xx = x * x, yy = y * y, zz = z * z;
  float xy = x * y, xz = x * z, yz = y * z;
  float wx = w * x, wy = w * y, wz = w * z;
  float* R = g->R;
  R[0] = 1.0f - 2.0f * (yy + zz); R[1] = 2.0f * (xy - wz);        R[2] = 2.0f * (xz + wy);
  R[3] = 2.0f * (xy + wz);        R[4] = 1.0f - 2.0f * (xx + zz); R[5] = 2.0f * (yz - wx);
  R[6] = 2.0f * (xz - wy);        R[7] = 2.0f * (yz + wx);        R[8] = 1.0f - 2.0f * (xx + yy);
  float* Mm = g->M;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) Mm[3 * i + j] = R[3 * i + j] * s[j];
#define SIGE(i, j) ((Mm[3 * (i) + 0] * Mm[3 * (j) + 0] + Mm[3 * (i) + 1] * Mm[3 * (j) + 1]) + Mm[3 * (i) + 2] * Mm[3 * (j) + 2])
  g->Sig[0] = SIGE(0, 0); g->Sig[1] = SIGE(0, 1); g->Sig[2] = SIGE(0, 2);
  g->Sig[3] = SIGE(1, 1); g->Sig[4] = SIGE(1, 2); g->Sig[5] = SIGE(2, 2);
#undef SIGE
  const float* S6 = g->Sig;
  /* v = Sigma jx, w = Sigma jy; cov2D = [jx.v, jx.w, jy.w] (projection.cu:102-114) */
  float v0 = (S6[0] * g->jx[0] + S6[1] * g->jx[1]) + S6[2] * g->jx[2];
  float v1 = (S6[1] * g->jx[0] + S6[3] * g->jx[1]) + S6[4] * g->jx[2];
  float v2 = (S6[2] * g->jx[0] + S6[4] * g->jx[1]) + S6[5] * g->jx[2];
  float w0 = (S6[0] * g->jy[0] + S6[1] * g->jy[1]) + S6[2] * g->jy[2];
  float w1 = (S6[1] * g->jy[0] + S6[3] * g->jy[1]) + S6[4] * g->jy[2];
  float w2 = (S6[2] * g->jy[0] + S6[4] * g->jy[1]) + S6[5] * g->jy[2];
  g->cxx = (g->jx[0] * v0 + g->jx[1] * v1) + g->jx[2] * v2;
  g->cxy = (g->jx[0] * w0 + g->jx[1] * w1) + g->jx[2] * w2;
  g->cyy = (g->jy[0] * w0 + g->jy[1] * w1) + g->jy[2] * w2;
  g->det = g->cxx * g->cyy - g->cxy * g->cxy;
  if (!(fabsf(g->det) >= 1e-8f)) return; /* projection.cu:115-119 */
  float dinv = 1.0f / g->det;
  g->ca = g->cyy * dinv;
  g->cb = -g->cxy * dinv;
  g->cc = g->cxx * dinv;

  /* clip / NDC / pixel (projection.cu:142-153, PD-1) */
  const float* F = c->proj;
  g->cx = ((F[0] * p[0] + F[4] * p[1]) + F[8] * p[2]) + F[12];
  g->cy = ((F[1] * p[0] + F[5] * p[1]) + F[9] * p[2]) + F[13];
  g->cz = ((F[2] * p[0] + F[6] * p[1]) + F[10] * p[2]) + F[14];
  g->cw = ((F[3] * p[0] + F[7] * p[1]) + F[11] * p[2]) + F[15];
  g->invW = 1.0f / (g->cw + 1e-5f);
  g->ndcx = g->cx * g->invW;
  g->ndcy = g->cy * g->invW;
  g->ndcz = g->cz * g->invW;
  g->px = ((g->ndcx + 1.0f) * 0.5f) * (float)c->W - 0.5f;
  g->py = ((g->ndcy + 1.0f) * 0.5f) * (float)c->H - 0.5f;
  g->visible = 1;
}

/* tile coordinate rule shared by count and emission (PD-3):
 * lo = clamp(trunc((x - r)/ts), 0, n), hi = clamp(trunc((x + r + ts-1)/ts), 0, n) */
static uint32_t tile_coord(float v, uint32_t n) {
  if (!(v > 0.0f)) return 0u;
  if (v >= (float)n) return n;
  return (uint32_t)v;
}

/* ------------------------------------------------------------------ */
/* stage 1: preprocess = projection + ADR culling + SH colour          */
/* projection.cu:7-154, ADRculling.cu:7-109, gaussianForwardPass.cu:172-264 */
/* ------------------------------------------------------------------ */
EXPORT int oracle_preprocess(int N, int M, int D, const float* means, const float* scales,
                             const float* rots, const float* opac, const float* sh,
                             const float* view, const float* proj, const float* campos,
                             const float* bg, int W, int H, int ts, float fx, float fy,
                             float* m2d, float* depth, float* conic, float* rgb, int32_t* radii,
                             uint32_t* rect, uint32_t* counts, float* cov3d) {
  OCam c;
  ocam_fill(&c, view, proj, campos, bg, W, H, ts, fx, fy);
  uint32_t tiles_x = (uint32_t)((W + ts - 1) / ts), tiles_y = (uint32_t)((H + ts - 1) / ts);
  float tsf = (float)ts, tsm1 = (float)(ts - 1);
  for (int i = 0; i < N; i++) {
    radii[i] = 0;
    counts[i] = 0;
    m2d[2 * i] = m2d[2 * i + 1] = 0.0f;
    depth[i] = 0.0f;
    conic[3 * i] = conic[3 * i + 1] = conic[3 * i + 2] = 0.0f;
    rect[4 * i] = rect[4 * i + 1] = rect[4 * i + 2] = rect[4 * i + 3] = 0;
    for (int k = 0; k < 6; k++) cov3d[6 * i + k] = 0.0f;
    const float* p = means + 3 * i;

    /* SH colour: direction = normalize(mean - campos), +0.5, unclamped (PD-8) */
    {
      float dx = p[0] - campos[0], dy = p[1] - campos[1], dz = p[2] - campos[2];
      float len = sqrtf((dx * dx + dy * dy) + dz * dz);
      float x = dx / len, y = dy / len, z = dz / len;
      float Y[16];
      sh_basis(D, x, y, z, Y);
      int nb = (D + 1) * (D + 1);
      const float* c0 = sh + (size_t)i * M * 3;
      for (int ch = 0; ch < 3; ch++) {
        float acc = Y[0] * c0[ch];
        for (int k = 1; k < nb; k++) acc = fmaf(Y[k], c0[3 * k + ch], acc);
        rgb[3 * i + ch] = acc + 0.5f;
      }
    }

    OGeom g;
    o_geometry(&c, p, scales + 3 * i, rots + 4 * i, &g);
    for (int k = 0; k < 6; k++) cov3d[6 * i + k] = g.Sig[k];
    if (!g.visible) continue;

    /* ADR radius (ADRculling.cu:39-76, PD-4, PD-5) */
    float half_tr = (g.cxx + g.cyy) * 0.5f;
    float ds = half_tr * half_tr - g.det;
    float root = sqrtf(fmaxf(0.1f, ds));
    float lmax = fmaxf(half_tr + root, half_tr - root);
    float r3 = 3.0f * sqrtf(lmax);
    float o = opac[i];
    if (!(o > 0.0039f)) continue;
    float lnr = w2w_logf(o * 255.0f);
    float radp = sqrtf((2.0f * lmax) * lnr); /* NaN when o < 1/255 -> fminf picks r3 */
    float r = fminf(radp, r3);
    if (!(r > 0.0f)) continue;
    /* rect (PD-3) */
    uint32_t x0 = tile_coord((g.px - r) / tsf, tiles_x);
    uint32_t y0 = tile_coord((g.py - r) / tsf, tiles_y);
    uint32_t x1 = tile_coord(((g.px + r) + tsm1) / tsf, tiles_x);
    uint32_t y1 = tile_coord(((g.py + r) + tsm1) / tsf, tiles_y);
    uint32_t cnt = (x1 > x0 && y1 > y0) ? (x1 - x0) * (y1 - y0) : 0u;
    radii[i] = (int32_t)ceilf(r);
    counts[i] = cnt;
    rect[4 * i + 0] = x0; rect[4 * i + 1] = y0; rect[4 * i + 2] = x1; rect[4 * i + 3] = y1;
    m2d[2 * i] = g.px;
    m2d[2 * i + 1] = g.py;
    depth[i] = g.ndcz;
    conic[3 * i] = g.ca; conic[3 * i + 1] = g.cb; conic[3 * i + 2] = g.cc;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* stage 2: inclusive scan of tile counts (gaussianRasterizer.cu:77-117) */
/* ------------------------------------------------------------------ */
EXPORT uint64_t oracle_inclusive_scan(int N, const uint32_t* in, uint32_t* out) {
  uint32_t acc = 0;
  for (int i = 0; i < N; i++) {
    acc += in[i];
    out[i] = acc;
  }
  return acc;
}

/* ------------------------------------------------------------------ */
/* stage 3: key emission (prepareSort.cu:178-205, 239-301, PD-3, PD-7)  */
/* key = tile << 32 | bits(depth), tiles row-major inside the rect.     */
/* ------------------------------------------------------------------ */
EXPORT int oracle_emit(int N, int W, int ts, const uint32_t* rect, const uint32_t* counts,
                       const uint32_t* incl_scan, const float* depth, uint64_t* keys,
                       uint32_t* vals) {
  uint32_t tiles_x = (uint32_t)((W + ts - 1) / ts);
  for (int i = 0; i < N; i++) {
    if (counts[i] == 0) continue;
    uint32_t off = (i == 0) ? 0u : incl_scan[i - 1];
    uint32_t db;
    memcpy(&db, depth + i, 4);
    for (uint32_t ty = rect[4 * i + 1]; ty < rect[4 * i + 3]; ty++)
      for (uint32_t tx = rect[4 * i + 0]; tx < rect[4 * i + 2]; tx++) {
        uint64_t tile = (uint64_t)(ty * tiles_x + tx);
        keys[off] = (tile << 32) | (uint64_t)db;
        vals[off] = (uint32_t)i;
        off++;
      }
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* stage 4: stable LSD radix sort of (u64 key, u32 value) over bits     */
/* [0, end_bit) -- the semantics of prepareSort.cu:303-360 (which runs  */
/* 2-bit digits over all 64 bits); 8-bit digits here, same result.      */
/* ------------------------------------------------------------------ */
EXPORT int oracle_sort_pairs(int64_t n, uint64_t* keys, uint32_t* vals, int end_bit) {
  if (n <= 1) return 0;
  uint64_t* k2 = (uint64_t*)malloc((size_t)n * 8);
  uint32_t* v2 = (uint32_t*)malloc((size_t)n * 4);
  if (!k2 || !v2) { free(k2); free(v2); return -1; }
  uint64_t *ks = keys, *kd = k2;
  uint32_t *vs = vals, *vd = v2;
  for (int shift = 0; shift < end_bit; shift += 8) {
    size_t cnt[257];
    memset(cnt, 0, sizeof(cnt));
    for (int64_t i = 0; i < n; i++) cnt[((ks[i] >> shift) & 0xffu) + 1]++;
    for (int d = 0; d < 256; d++) cnt[d + 1] += cnt[d];
    for (int64_t i = 0; i < n; i++) {
      size_t pos = cnt[(ks[i] >> shift) & 0xffu]++;
      kd[pos] = ks[i];
      vd[pos] = vs[i];
    }
    uint64_t* tk = ks; ks = kd; kd = tk;
    uint32_t* tv = vs; vs = vd; vd = tv;
  }
  if (ks != keys) {
    memcpy(keys, ks, (size_t)n * 8);
    memcpy(vals, vs, (size_t)n * 4);
  }
  free(k2);
  free(v2);
  return 0;
}

/* ------------------------------------------------------------------ */
/* stage 5: per-tile [start,end) (gaussianTiles.cu:6-33); empty = 0,0    */
/* ------------------------------------------------------------------ */
EXPORT int oracle_ranges(int64_t K, const uint64_t* keys, int T, uint32_t* ranges) {
  memset(ranges, 0, (size_t)T * 8);
  for (int64_t i = 0; i < K; i++) {
    uint32_t t = (uint32_t)(keys[i] >> 32);
    if (t >= (uint32_t)T) return -1;
    if (i == 0 || (uint32_t)(keys[i - 1] >> 32) != t) ranges[2 * t] = (uint32_t)i;
    if (i == K - 1 || (uint32_t)(keys[i + 1] >> 32) != t) ranges[2 * t + 1] = (uint32_t)(i + 1);
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* stage 6: front-to-back blend (pixelGaussiansFull.cu:3-98,             */
/* pixelGaussiansDev.cuh:5-96, PD-9).  Pixel centres at integer coords. */
/* alpha = min(0.99, o G), skip alpha < 1/255, stop when T(1-alpha) <    */
/* 1e-4 (that instance excluded).  n_contrib = 1-based last contributor. */
/* ------------------------------------------------------------------ */
static inline float o_power(float ca, float cb, float cc, float dx, float dy) {
  /* -0.5 (a dx^2 + c dy^2) - b dx dy, fixed order */
  float q = fmaf(ca * dx, dx, (cc * dy) * dy);
  return fmaf(-0.5f, q, -((cb * dx) * dy));
}

EXPORT int oracle_blend(int W, int H, int ts, const uint32_t* ranges, const uint32_t* ids,
                        const float* m2d, const float* conic, const float* opac,
                        const float* rgb, const float* depth, const float* bg, float* out_color,
                        float* out_alpha, float* out_depth, uint32_t* n_contrib, float* final_T,
                        int max_rec, uint32_t* rec_count, uint32_t* rec_ids) {
  int tiles_x = (W + ts - 1) / ts, tiles_y = (H + ts - 1) / ts;
  const float a_min = 0x1.010102p-8f; /* 1/255 */
  for (int ty = 0; ty < tiles_y; ty++)
    for (int tx = 0; tx < tiles_x; tx++) {
      uint32_t s = ranges[2 * (ty * tiles_x + tx)], e = ranges[2 * (ty * tiles_x + tx) + 1];
      for (int py = ty * ts; py < ty * ts + ts && py < H; py++)
        for (int px = tx * ts; px < tx * ts + ts && px < W; px++) {
          size_t pid = (size_t)py * W + px;
          float T = 1.0f, C0 = 0.0f, C1 = 0.0f, C2 = 0.0f, A = 0.0f, Dd = 0.0f;
          uint32_t last = 0, nrec = 0;
          for (uint32_t k = s; k < e; k++) {
            uint32_t g = ids[k];
            float dx = m2d[2 * g] - (float)px, dy = m2d[2 * g + 1] - (float)py;
            float pw = o_power(conic[3 * g], conic[3 * g + 1], conic[3 * g + 2], dx, dy);
            if (pw > 0.0f) continue;
            float G = w2w_expf(pw);
            float a = fminf(0.99f, opac[g] * G);
            if (a < a_min) continue;
            float nT = T * (1.0f - a);
            if (nT < 1e-4f) break;
            float aT = a * T;
            C0 = fmaf(rgb[3 * g], aT, C0);
            C1 = fmaf(rgb[3 * g + 1], aT, C1);
            C2 = fmaf(rgb[3 * g + 2], aT, C2);
            A = A + aT;
            Dd = fmaf(depth[g], aT, Dd);
            T = nT;
            last = k - s + 1;
            if (rec_ids) {
              if ((int)nrec >= max_rec) return -2;
              rec_ids[pid * (size_t)max_rec + nrec] = g;
            }
            nrec++;
          }
          out_color[pid] = fmaf(T, bg[0], C0);
          out_color[(size_t)W * H + pid] = fmaf(T, bg[1], C1);
          out_color[2 * (size_t)W * H + pid] = fmaf(T, bg[2], C2);
          out_alpha[pid] = A;
          out_depth[pid] = Dd;
          n_contrib[pid] = last;
          final_T[pid] = T;
          if (rec_count) rec_count[pid] = nrec;
        }
    }
  return 0;
}

/* ------------------------------------------------------------------ */
/* backward stage 1: reverse blend (backward.cu:521-665 structure).     */
/* Starts from the stored final T and n_contrib; T <- T/(1-alpha).      */
/* Outputs per Gaussian: dL/dmean2D in NDC units (x 0.5W, 0.5H as       */
/* backward.cu:586-587), dL/dconic (a,b,c as partials of the power      */
/* form), dL/dopacity, dL/drgb, dL/ddepth.  The 0.99 clamp is honoured  */
/* (PD-10): a clamped alpha passes no gradient to o or G.               */
/* ------------------------------------------------------------------ */
EXPORT int oracle_blend_backward(int W, int H, int ts, const uint32_t* ranges,
                                 const uint32_t* ids, const float* m2d, const float* conic,
                                 const float* opac, const float* rgb, const float* depth,
                                 const float* bg, const uint32_t* n_contrib, const float* final_T,
                                 const float* dL_dcolor, const float* dL_dalpha,
                                 const float* dL_ddepth_img, float* g_m2d, float* g_conic,
                                 float* g_opac, float* g_rgb, float* g_depth) {
  int tiles_x = (W + ts - 1) / ts, tiles_y = (H + ts - 1) / ts;
  const float a_min = 0x1.010102p-8f;
  float hw = 0.5f * (float)W, hh = 0.5f * (float)H;
  size_t HW = (size_t)W * H;
  for (int ty = 0; ty < tiles_y; ty++)
    for (int tx = 0; tx < tiles_x; tx++) {
      uint32_t s = ranges[2 * (ty * tiles_x + tx)];
      for (int py = ty * ts; py < ty * ts + ts && py < H; py++)
        for (int px = tx * ts; px < tx * ts + ts && px < W; px++) {
          size_t pid = (size_t)py * W + px;
          uint32_t last = n_contrib[pid];
          if (last == 0) continue;
          float Tf = final_T[pid];
          float T = Tf;
          float gC0 = dL_dcolor[pid], gC1 = dL_dcolor[HW + pid], gC2 = dL_dcolor[2 * HW + pid];
          float gA = dL_dalpha[pid], gD = dL_ddepth_img[pid];
          float bgdot = (bg[0] * gC0 + bg[1] * gC1) + bg[2] * gC2;
          float acc0 = 0.0f, acc1 = 0.0f, acc2 = 0.0f, accD = 0.0f, accA = 0.0f;
          float la = 0.0f, lc0 = 0.0f, lc1 = 0.0f, lc2 = 0.0f, ld = 0.0f;
          for (int64_t k = (int64_t)s + last - 1; k >= (int64_t)s; k--) {
            uint32_t g = ids[k];
            float dx = m2d[2 * g] - (float)px, dy = m2d[2 * g + 1] - (float)py;
            float ca = conic[3 * g], cb = conic[3 * g + 1], cc = conic[3 * g + 2];
            float pw = o_power(ca, cb, cc, dx, dy);
            if (pw > 0.0f) continue;
            float G = w2w_expf(pw);
            float o = opac[g];
            float oG = o * G;
            float a = fminf(0.99f, oG);
            if (a < a_min) continue;
            T = T / (1.0f - a);
            float aT = a * T;
            float c0 = rgb[3 * g], c1 = rgb[3 * g + 1], c2 = rgb[3 * g + 2], dd = depth[g];
            acc0 = la * lc0 + (1.0f - la) * acc0;
            acc1 = la * lc1 + (1.0f - la) * acc1;
            acc2 = la * lc2 + (1.0f - la) * acc2;
            accD = la * ld + (1.0f - la) * accD;
            accA = la + (1.0f - la) * accA;
            float dLda = ((c0 - acc0) * gC0 + (c1 - acc1) * gC1) + (c2 - acc2) * gC2;
            dLda = dLda + (dd - accD) * gD;
            dLda = dLda + (1.0f - accA) * gA;
            dLda = dLda * T;
            dLda = dLda - (Tf / (1.0f - a)) * bgdot;
            lc0 = c0; lc1 = c1; lc2 = c2; ld = dd; la = a;
            g_rgb[3 * g] += aT * gC0;
            g_rgb[3 * g + 1] += aT * gC1;
            g_rgb[3 * g + 2] += aT * gC2;
            g_depth[g] += aT * gD;
            if (oG <= 0.99f) {
              g_opac[g] += G * dLda;
              float dLdp = (o * dLda) * G; /* dL/dpower */
              g_conic[3 * g] += (-0.5f * dx * dx) * dLdp;
              g_conic[3 * g + 1] += (-dx * dy) * dLdp;
              g_conic[3 * g + 2] += (-0.5f * dy * dy) * dLdp;
              g_m2d[2 * g] += (-(ca * dx + cb * dy) * dLdp) * hw;
              g_m2d[2 * g + 1] += (-(cb * dx + cc * dy) * dLdp) * hh;
            }
          }
        }
    }
  return 0;
}

/* ------------------------------------------------------------------ */
/* backward stage 2: per-Gaussian chain rule (the role of               */
/* backward.cu:27-227 computeCov2DCUDA and :229-519                     */
/* backwardPreprocessingKernel), differentiating the forward above      */
/* exactly (PD-10: no 0.3 dilation, no 1.3 tan clamp, degree-aware SH,  */
/* coefficient 6 = 0.3, no clamp mask, NDC depth).                      */
/* ------------------------------------------------------------------ */
EXPORT int oracle_preprocess_backward(int N, int M, int D, const float* means, const float* scales,
                                      const float* rots, const float* sh, const float* view,
                                      const float* proj, const float* campos, int W, int H,
                                      float fx, float fy, const int32_t* radii,
                                      const float* g_m2d, const float* g_conic,
                                      const float* g_rgb, const float* g_depth,
                                      float* dL_dmeans, float* dL_dcov, float* dL_dsh,
                                      float* dL_dscales, float* dL_drots) {
  OCam c;
  float zero3[3] = {0, 0, 0};
  ocam_fill(&c, view, proj, campos, zero3, W, H, 16, fx, fy);
  const float* V = view;
  const float* F = proj;
  float limx = (float)W / (2.0f * fx), limy = (float)H / (2.0f * fy);
  for (int i = 0; i < N; i++) {
    float* gm = dL_dmeans + 3 * i;
    gm[0] = gm[1] = gm[2] = 0.0f;
    for (int k = 0; k < 6; k++) dL_dcov[6 * i + k] = 0.0f;
    for (int k = 0; k < M * 3; k++) dL_dsh[(size_t)i * M * 3 + k] = 0.0f;
    for (int k = 0; k < 3; k++) dL_dscales[3 * i + k] = 0.0f;
    for (int k = 0; k < 4; k++) dL_drots[4 * i + k] = 0.0f;
    if (radii[i] <= 0) continue;
    const float* p = means + 3 * i;
    const float* s = scales + 3 * i;
    const float* q = rots + 4 * i;
    OGeom g;
    o_geometry(&c, p, s, q, &g);
    if (!g.visible) continue;

    /* --- colour: SH coefficients and direction --- */
    {
      float dx = p[0] - campos[0], dy = p[1] - campos[1], dz = p[2] - campos[2];
      float len = sqrtf((dx * dx + dy * dy) + dz * dz);
      float x = dx / len, y = dy / len, z = dz / len;
      float Y[16], gx[16], gy[16], gz[16];
      sh_basis(D, x, y, z, Y);
      sh_basis_grad(D, x, y, z, gx, gy, gz);
      int nb = (D + 1) * (D + 1);
      const float* cf = sh + (size_t)i * M * 3;
      const float* gr = g_rgb + 3 * i;
      float gdx = 0.0f, gdy = 0.0f, gdz = 0.0f;
      for (int k = 0; k < nb; k++) {
        for (int ch = 0; ch < 3; ch++) dL_dsh[(size_t)i * M * 3 + 3 * k + ch] = Y[k] * gr[ch];
        float dot = (cf[3 * k] * gr[0] + cf[3 * k + 1] * gr[1]) + cf[3 * k + 2] * gr[2];
        gdx = fmaf(gx[k], dot, gdx);
        gdy = fmaf(gy[k], dot, gdy);
        gdz = fmaf(gz[k], dot, gdz);
      }
      float proj_d = (gdx * x + gdy * y) + gdz * z;
      gm[0] += (gdx - proj_d * x) / len;
      gm[1] += (gdy - proj_d * y) / len;
      gm[2] += (gdz - proj_d * z) / len;
    }

    /* --- NDC mean and depth through the full projection --- */
    {
      float gnx = g_m2d[2 * i], gny = g_m2d[2 * i + 1], gnz = g_depth[i];
      float gcx = gnx * g.invW, gcy = gny * g.invW, gcz = gnz * g.invW;
      float ginvW = (gnx * g.cx + gny * g.cy) + gnz * g.cz;
      float gcw = -ginvW * g.invW * g.invW;
      for (int j = 0; j < 3; j++)
        gm[j] += ((F[4 * j + 0] * gcx + F[4 * j + 1] * gcy) + F[4 * j + 2] * gcz) + F[4 * j + 3] * gcw;
    }

    /* --- conic -> cov2D --- */
    float ga = g_conic[3 * i], gb = g_conic[3 * i + 1], gc = g_conic[3 * i + 2];
    float a = g.ca, b = g.cb, cq = g.cc, h = 0.5f * gb;
    float q00 = a * ga + b * h, q01 = a * h + b * gc, q10 = b * ga + cq * h, q11 = b * h + cq * gc;
    float g_cxx = -(q00 * a + q01 * b);
    float g_cxy = -2.0f * (q00 * b + q01 * cq);
    float g_cyy = -(q10 * b + q11 * cq);

    /* --- cov2D -> Sigma and J --- */
    const float* S6 = g.Sig;
    float v[3], w[3];
    v[0] = (S6[0] * g.jx[0] + S6[1] * g.jx[1]) + S6[2] * g.jx[2];
    v[1] = (S6[1] * g.jx[0] + S6[3] * g.jx[1]) + S6[4] * g.jx[2];
    v[2] = (S6[2] * g.jx[0] + S6[4] * g.jx[1]) + S6[5] * g.jx[2];
    w[0] = (S6[0] * g.jy[0] + S6[1] * g.jy[1]) + S6[2] * g.jy[2];
    w[1] = (S6[1] * g.jy[0] + S6[3] * g.jy[1]) + S6[4] * g.jy[2];
    w[2] = (S6[2] * g.jy[0] + S6[4] * g.jy[1]) + S6[5] * g.jy[2];
    float hxy = 0.5f * g_cxy;
    float GS[9]; /* symmetric dL/dSigma = J^T G2 J */
    for (int k = 0; k < 3; k++)
      for (int l = 0; l < 3; l++)
        GS[3 * k + l] = (g_cxx * g.jx[k] * g.jx[l] + hxy * (g.jx[k] * g.jy[l] + g.jy[k] * g.jx[l])) +
                        g_cyy * g.jy[k] * g.jy[l];
    float* dc = dL_dcov + 6 * i;
    dc[0] = GS[0]; dc[1] = 2.0f * GS[1]; dc[2] = 2.0f * GS[2];
    dc[3] = GS[4]; dc[4] = 2.0f * GS[5]; dc[5] = GS[8];
    float gjx[3], gjy[3];
    for (int k = 0; k < 3; k++) {
      gjx[k] = 2.0f * g_cxx * v[k] + g_cxy * w[k];
      gjy[k] = g_cxy * v[k] + 2.0f * g_cyy * w[k];
    }

    /* --- J -> view-space point --- */
    float zi = 1.0f / g.vz, zi2 = zi * zi;
    float g_zi = 0.0f, g_zi2 = 0.0f, g_txc = 0.0f, g_tyc = 0.0f;
    for (int k = 0; k < 3; k++) {
      float right = V[4 * k + 0], up = V[4 * k + 1], fwd = V[4 * k + 2];
      g_zi += gjx[k] * fx * right + gjy[k] * fy * up;
      g_zi2 -= gjx[k] * fx * g.txc * fwd + gjy[k] * fy * g.tyc * fwd;
      g_txc -= gjx[k] * fx * fwd * zi2;
      g_tyc -= gjy[k] * fy * fwd * zi2;
    }
    float g_vz = g_zi * (-zi * zi) + g_zi2 * (-2.0f * zi2 * zi);
    /* txc = rxc * vz */
    float g_rxc = g_txc * g.vz, g_ryc = g_tyc * g.vz;
    g_vz += g_txc * g.rxc + g_tyc * g.ryc;
    float g_rx, g_ry;
    if (g.clamped) {
      float s3 = g.s * g.s * g.s;
      float lx2 = limx * limx, ly2 = limy * limy;
      g_rx = g_rxc * (g.s - s3 * g.rx * g.rx / lx2) + g_ryc * (-s3 * g.rx * g.ry / lx2);
      g_ry = g_rxc * (-s3 * g.rx * g.ry / ly2) + g_ryc * (g.s - s3 * g.ry * g.ry / ly2);
    } else {
      g_rx = g_rxc;
      g_ry = g_ryc;
    }
    float g_vx = g_rx / g.vz, g_vy = g_ry / g.vz;
    g_vz -= (g_rx * g.rx + g_ry * g.ry) / g.vz;
    /* v = t / tw */
    float g_tx = g_vx / g.tw, g_ty = g_vy / g.tw, g_tz = g_vz / g.tw;
    float g_tw = -((g_vx * g.vx + g_vy * g.vy) + g_vz * g.vz) / g.tw;
    for (int j = 0; j < 3; j++)
      gm[j] += ((V[4 * j + 0] * g_tx + V[4 * j + 1] * g_ty) + V[4 * j + 2] * g_tz) + V[4 * j + 3] * g_tw;

    /* --- Sigma -> scales and quaternion --- */
    const float* Mm = g.M;
    const float* R = g.R;
    float gM[9];
    for (int r = 0; r < 3; r++)
      for (int col = 0; col < 3; col++)
        gM[3 * r + col] = 2.0f * ((GS[3 * r + 0] * Mm[col] + GS[3 * r + 1] * Mm[3 + col]) + GS[3 * r + 2] * Mm[6 + col]);
    float gR[9];
    for (int r = 0; r < 3; r++)
      for (int col = 0; col < 3; col++) gR[3 * r + col] = gM[3 * r + col] * s[col];
    for (int col = 0; col < 3; col++)
      dL_dscales[3 * i + col] = (gM[col] * R[col] + gM[3 + col] * R[3 + col]) + gM[6 + col] * R[6 + col];
    float qw = q[0], qx = q[1], qy = q[2], qz = q[3];
    float* gq = dL_drots + 4 * i;
    gq[0] = 2.0f * (((-qz * gR[1] + qy * gR[2]) + (qz * gR[3] - qx * gR[5])) + (-qy * gR[6] + qx * gR[7]));
    gq[1] = 2.0f * ((((qy * gR[1] + qz * gR[2]) + (qy * gR[3] - 2.0f * qx * gR[4])) + (-qw * gR[5] + qz * gR[6])) + (qw * gR[7] - 2.0f * qx * gR[8]));
    gq[2] = 2.0f * ((((-2.0f * qy * gR[0] + qx * gR[1]) + (qw * gR[2] + qx * gR[3])) + (qz * gR[5] - qw * gR[6])) + (qz * gR[7] - 2.0f * qy * gR[8]));
    gq[3] = 2.0f * ((((-2.0f * qz * gR[0] - qw * gR[1]) + (qx * gR[2] + qw * gR[3])) + (-2.0f * qz * gR[4] + qy * gR[5])) + (qx * gR[6] + qy * gR[7]));
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* occupancy field for mesh export                                     */
/* ------------------------------------------------------------------ */

/* compute_3d_gaussian_coefficient_device (compute_voxel_field.cu:6-23), with
 * libm expf standing for CUDA's expf. */
static float vox_coef(float x, float y, float z, float a, float b, float c, float d, float e, float f) {
    float det = a * d * f + 2 * e * c * b - e * e * a - c * c * d - b * b * f;
    float inv_det = 1.0f / (det + 1e-24f);
    float inv_a = (d * f - e * e) * inv_det;
    float inv_b = (e * c - b * f) * inv_det;
    float inv_c = (e * b - c * d) * inv_det;
    float inv_d = (a * f - c * c) * inv_det;
    float inv_e = (b * c - e * a) * inv_det;
    float inv_f = (a * d - b * b) * inv_det;
    float power = -0.5f * (x * x * inv_a + y * y * inv_d + z * z * inv_f) - x * y * inv_b - x * z * inv_c -
                  y * z * inv_e;
    if (power > 0) power = -1e10f;
    return expf(power);
}

/* compute_voxel_field_kernel (compute_voxel_field.cu:25-79), one grid point
 * at a time: every Gaussian in index order, kept when its mean lies in the
 * relaxed box of the point's block; occ[ix + iy res + iz res^2]. */
EXPORT void oracle_voxel_field(float* occ, const float* means, const float* covs, const float* opac, int n,
                               int res, int num_blocks, float relax_ratio) {
    for (int idz = 0; idz < res; idz++)
        for (int idy = 0; idy < res; idy++)
            for (int idx = 0; idx < res; idx++) {
                float x = -1.0f + 2.0f * idx / (res - 1);
                float y = -1.0f + 2.0f * idy / (res - 1);
                float z = -1.0f + 2.0f * idz / (res - 1);
                float block_size = 2.0f / num_blocks;
                int block_x = (int)((x + 1.0f) / block_size);
                int block_y = (int)((y + 1.0f) / block_size);
                int block_z = (int)((z + 1.0f) / block_size);
                float vmin_x = -1.0f + block_x * block_size - block_size * relax_ratio;
                float vmax_x = -1.0f + (block_x + 1) * block_size + block_size * relax_ratio;
                float vmin_y = -1.0f + block_y * block_size - block_size * relax_ratio;
                float vmax_y = -1.0f + (block_y + 1) * block_size + block_size * relax_ratio;
                float vmin_z = -1.0f + block_z * block_size - block_size * relax_ratio;
                float vmax_z = -1.0f + (block_z + 1) * block_size + block_size * relax_ratio;
                float val = 0.0f;
                for (int g = 0; g < n; g++) {
                    float mx = means[g * 3 + 0], my = means[g * 3 + 1], mz = means[g * 3 + 2];
                    if (mx < vmin_x || mx > vmax_x || my < vmin_y || my > vmax_y || mz < vmin_z || mz > vmax_z)
                        continue;
                    const float* cv = covs + 6 * g;
                    float w = vox_coef(x - mx, y - my, z - mz, cv[0], cv[1], cv[2], cv[3], cv[4], cv[5]);
                    val += opac[g] * w;
                }
                occ[idx + idy * res + idz * res * res] = val;
            }
}
