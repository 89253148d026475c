// Sampling and shading of the path estimator (device side).
//
// The renderer follows DumbRenderer::PixelRenderProcess (DumbRenderer.cu:174-443):
// camera samples per subpixel, one Octree::cast per layer with the previous
// layer's triangle excluded, light photons cast from the light toward the hit
// point and counted when they first hit the same triangle or land within
// sqrt(8 eps) of it, and the layer colour reflected up the path.  The
// materials and the light are the plugins the benchmark scenes need (Lambert,
// perfect mirror, Schlick-Fresnel metal, one-sided area quad light, Russian
// roulette); random numbers are counter-based Philox4x32-10 keyed by
// (seed) with counter (pixel, sample, layer << 8 | purpose, request), so any
// schedule draws the same numbers for the same sample.  Every fp32 expression
// here is evaluated in a fixed order without contraction, so rays, hits and
// sample directions are bit-identical to a host evaluation of the same
// expressions.
#pragma once
#include "dr_geom.h"
#include "../../include/dumbray.h"

namespace dr {

constexpr float kInvPi = 0.31830988618379067154f;
constexpr int kMaxLights = 16;

DR_HD void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                         uint32_t out[4]) {
  for (int r = 0; r < 10; r++) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0, hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Four uniforms in [0,1) (24-bit) for one (pixel, sample, layer, purpose).
// purpose 0 = lens jitter, 1 = BSDF direction (u0, u1) + Russian roulette (u2), 16 + l = light l.
DR_HD void rng4(const dr_render_params &P, uint32_t pix, uint32_t samp, uint32_t layer, uint32_t purpose, float u[4]) {
  uint32_t o[4];
  philox4x32_10(pix, samp, (layer << 8) | purpose, P.request, P.seed_lo, P.seed_hi, o);
  for (int k = 0; k < 4; k++) u[k] = (float)(o[k] >> 8) * (1.0f / 16777216.0f);
}

// sin and cos of 2*pi*u, u in [0,1): quadrant from 4u, Taylor series on [0, pi/2).
DR_HD void sincos_2pi(float u, float &s, float &c) {
  const float x = u * 4.0f;
  const int q = (int)x;
  const float f = x - (float)q;
  const float a = f * 1.57079637f;
  const float a2 = a * a;
  const float sa = a * (1.0f + a2 * (-0.166666667f + a2 * (0.00833333377f + a2 * (-0.000198412701f +
                   a2 * (2.75573188e-06f + a2 * (-2.50521079e-08f))))));
  const float ca = 1.0f + a2 * (-0.5f + a2 * (0.0416666679f + a2 * (-0.00138888892f + a2 * (2.48015876e-05f +
                   a2 * (-2.75573188e-07f + a2 * 2.08767570e-09f)))));
  switch (q & 3) {
    case 0: s = sa; c = ca; break;
    case 1: s = ca; c = -sa; break;
    case 2: s = -sa; c = -ca; break;
    default: s = -ca; c = sa; break;
  }
}

DR_HD V3 normalized(V3 a) {
  const float il = 1.0f / __builtin_sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);   // Vector3::normalized
  return a * il;
}

// sin and cos of an angle in radians: sincos_2pi of the angle in turns, reduced to [0, 1).
DR_HD void sincos_rad(float a, float &s, float &c) {
  float t = a * 0.159154943f;                 // 1 / (2 pi)
  t = t - __builtin_floorf(t);
  sincos_2pi(t, s, c);
}

// arcsin on [0, 1]: the Taylor series of arcsin (terms to x^21) for x <= 1/2, and
// asin(x) = pi/2 - 2 asin(sqrt((1 - x) / 2)) above.
DR_HD float asin_series(float x) {
  const float z = x * x;
  float p = 7.312525874e-03f;
  p = p * z + 8.390335810e-03f;
  p = p * z + 9.761609529e-03f;
  p = p * z + 1.155180090e-02f;
  p = p * z + 1.396484375e-02f;
  p = p * z + 1.735276442e-02f;
  p = p * z + 2.237215909e-02f;
  p = p * z + 3.038194444e-02f;
  p = p * z + 4.464285714e-02f;
  p = p * z + 7.500000000e-02f;
  p = p * z + 1.666666667e-01f;
  return x + x * (z * p);
}
DR_HD float asin_unit(float x) {
  if (x <= 0.5f) return asin_series(x);
  return 1.57079637f - 2.0f * asin_series(__builtin_sqrtf((1.0f - x) * 0.5f));
}

// Pixel size of a boxing mode (DumbRenderer.cu:190-200).
DR_HD float pixel_size(float W, float H, int boxing) {
  if (boxing == DR_BOXING_WIDTH) return 1.0f / W;
  if (boxing == DR_BOXING_MIN) return 1.0f / ((H <= W) ? H : W);
  if (boxing == DR_BOXING_MAX) return 1.0f / ((H >= W) ? H : W);
  return 1.0f / H;
}

// Division of a non-negative int by a divisor d >= 1 through its reciprocal
// m = floor((2^32 - 1) / d): for n < 2^32, n * m / 2^32 lies in (n / d - 1, n / d], so the
// high word of n * m is n / d or one less, and one comparison corrects it.
struct UDiv { uint32_t d, m; };
DR_HD UDiv make_udiv(uint32_t d) { return {d, 0xffffffffu / d}; }
DR_HD uint32_t udiv(uint32_t n, UDiv D) {
  const uint32_t q = (uint32_t)(((uint64_t)n * D.m) >> 32);
  return (n - q * D.d >= D.d) ? q + 1 : q;
}

// Values of one request's camera, framebuffer mapping and Russian roulette that depend on
// the parameters alone.  The host forms them (frame_k, the same fp32 expressions the kernels
// would evaluate) and passes them to the kernels, where they sit in scalar registers; formed
// inside the path loop, the compiler hoists them out of it into vector registers.
struct FrameK {
  UDiv nbw;              // framebuffer blocks per row
  UDiv fy;               // fsaa_y: sample s is subpixel (s / fsaa_y, s % fsaa_y)
  float hw, hh;          // W / 2, H / 2
  float ps, pdx, pdy;    // pixel size; subpixel width and height
  float jx, ox, jy, oy;  // lens jitter: range offx - (-offx) and start -offx (x), the same for y
  float zc[4];           // cam[12] * cam[9], cam[10], cam[11], cam[12]
  float amount, keep;    // blend weight of this iteration (1 / iteration) and 1 - amount
  float inv_rr;          // 1 / rr_survive: the weight of a path that survives Russian roulette
};
DR_HD FrameK frame_k(const dr_render_params &P) {
  FrameK K;
  const float W = (float)P.width, H = (float)P.height;
  K.nbw = make_udiv((uint32_t)((P.width + 8 - 1) / 8));
  K.fy = make_udiv((uint32_t)P.fsaa_y);
  K.hw = W / 2.0f;
  K.hh = H / 2.0f;
  K.ps = pixel_size(W, H, P.boxing);
  K.pdx = K.ps / ((float)P.fsaa_x);
  K.pdy = K.ps / ((float)P.fsaa_y);
  const float offx = K.pdx * 0.5f, offy = K.pdy * 0.5f;
  K.jx = offx - (-offx);
  K.ox = -offx;
  K.jy = offy - (-offy);
  K.oy = -offy;
  for (int k = 0; k < 4; k++) K.zc[k] = P.cam[12] * P.cam[9 + k];
  K.amount = (P.iteration <= 1) ? 1.0f : (1.0f / ((float)P.iteration));
  K.keep = 1.0f - K.amount;
  K.inv_rr = 1.0f / P.rr_survive;
  return K;
}

// Camera sample of subpixel (i, j) of pixel (px, py): DumbRenderer.cu:190-232 (pixel
// size by boxing mode), SimpleStochasticLense jitter (SimpleStochasticLense.impl.cuh:11-21 with
// DumbRand::range = u*(max-min)+min) or DefaultPerspectiveLense, Camera transform
// (Transform.impl.h:166-168).  K = frame_k(P).
DR_HD void camera_ray(const dr_render_params &P, const FrameK &K, int px, int py, int i, int j, float u0, float u1,
                      float u2, float u3, V3 &o, V3 &d) {
  const float ppx = ((float)px - K.hw) * K.ps;
  const float ppy = (K.hh - (float)py) * K.ps;
  const float sx = ppx + (((float)i) + 0.5f) * K.pdx;
  const float sy = ppy + (((float)j) + 0.5f) * K.pdy;
  const float *C = P.cam;
  if (P.lens == DR_SPHERICAL_SEGMENT_LENS) {
    // Lenses::SphericalSegmentLense (SphericalSegmentLense.impl.cuh:13-30): jittered angles
    // alpha * position / RADIAN, target = direction * distance, origin on a disc of radius
    // asin(u) * sensor; the origin goes through the camera's point map, the unit direction
    // (target - origin) through Transform::direction.
    const float ha = ((sx + ((u0 * K.jx) + K.ox)) * P.lens_p[0]) / 57.29577951308233f;
    const float va = ((sy + ((u1 * K.jy) + K.oy)) * P.lens_p[1]) / 57.29577951308233f;
    float hs, hc, vs, vc;
    sincos_rad(ha, hs, hc);
    sincos_rad(va, vs, vc);
    const V3 target = v3(hs * vc, vs, hc * vc) * P.lens_p[2];
    float rs, rc;
    sincos_2pi(u2, rs, rc);                     // range(0, 2*PI)
    const V3 org = (v3(rc, rs, 0.0f) * asin_unit(u3)) * P.lens_p[3];
    const V3 dl = normalized(target - org);
    const float *R = P.cam_rows;
    o = v3(((org.x * R[0] + org.y * R[1]) + org.z * R[2]) + C[0], ((org.x * R[3] + org.y * R[4]) + org.z * R[5]) + C[1],
           ((org.x * R[6] + org.y * R[7]) + org.z * R[8]) + C[2]);
    d = v3((dl.x * C[3] + dl.y * C[6]) + dl.z * C[9], (dl.x * C[4] + dl.y * C[7]) + dl.z * C[10],
           (dl.x * C[5] + dl.y * C[8]) + dl.z * C[11]);
    return;
  }
  o = v3(C[0], C[1], C[2]);
  if (P.lens == DR_PERSPECTIVE_LENS) {
    // DefaultPerspectiveLense (DefaultPerspectiveLense.impl.cuh:12-15): subpixel centre,
    // direction normalized in lens space before the camera transform (Vector3::normalized
    // of (sx, sy, cam[12]))
    const float il = 1.0f / __builtin_sqrtf(sx * sx + sy * sy + K.zc[3]);
    const float lx = sx * il, ly = sy * il, lz = C[12] * il;
    d = v3((lx * C[3] + ly * C[6]) + lz * C[9], (lx * C[4] + ly * C[7]) + lz * C[10],
           (lx * C[5] + ly * C[8]) + lz * C[11]);
  } else {
    const float lx = sx + ((u0 * K.jx) + K.ox);
    const float ly = sy + ((u1 * K.jy) + K.oy);
    d = v3((lx * C[3] + ly * C[6]) + K.zc[0], (lx * C[4] + ly * C[7]) + K.zc[1], (lx * C[5] + ly * C[8]) + K.zc[2]);
  }
}

// Unit geometric normal n = (b-a)x(c-a) normalised, turned to face the incoming ray.
DR_HD V3 facing_normal(V3 n, V3 du) {
  V3 nu = normalized(n);
  if (dot(nu, du) > 0) nu = v3(-nu.x, -nu.y, -nu.z);
  return nu;
}

// Cosine-weighted direction about unit n (orthonormal basis of Duff et al., JCGT 2017).
DR_HD V3 cosine_dir(V3 n, float u1, float u2) {
  const float sign = __builtin_copysignf(1.0f, n.z);
  const float a = -1.0f / (sign + n.z);
  const float b = n.x * n.y * a;
  const V3 b1 = v3(1.0f + ((sign * n.x) * n.x) * a, sign * b, -sign * n.x);
  const V3 b2 = v3(b, sign + (n.y * n.y) * a, -n.y);
  const float r = __builtin_sqrtf(u1);
  float s, c;
  sincos_2pi(u2, s, c);
  const float x = r * c, y = r * s, z = __builtin_sqrtf(1.0f - u1);
  return ((b1 * x) + (b2 * y)) + (n * z);
}

// Vector3::reflectionOnUnitVector (Vector3.impl.h:393-396)
DR_HD V3 reflect_unit(V3 v, V3 n) {
  const float mul = (2.0f * dot(v, n));
  return v3(v.x - (n.x * mul), v.y - (n.y * mul), v.z - (n.z * mul));
}

// Sampling decision at `layer` (requestIndirectSamples + Russian roulette); K = frame_k(P).
DR_HD bool next_sample(const dr_render_params &P, const FrameK &K, int kind, uint32_t pix, uint32_t samp, int layer,
                       V3 nu, V3 du, V3 &dir, float &weight) {
  if (layer >= P.max_bounces) return false;
  float u[4];
  rng4(P, pix, samp, (uint32_t)layer, 1, u);
  weight = 1.0f;
  if (layer >= P.rr_start) {
    if (u[2] >= P.rr_survive) return false;
    weight = K.inv_rr;
  }
  dir = (kind == DR_LAMBERT) ? cosine_dir(nu, u[0], u[1]) : reflect_unit(du, nu);
  return true;
}

// Schlick factor of the metal for the reflected unit direction (DumbTools::fresnel, DumbStructs.cuh:136-140).
DR_HD float schlick5(V3 nu, V3 dir) {
  const float val = (1.0f - dot(nu, dir));
  const float sq = (val * val);
  return (sq * sq) * val;
}

// Area-quad light photon toward p: shadow ray (q, w), cosine at p and cos_l * A / d^2.
DR_HD bool light_photon(const dr_render_params &P, const dr_light &L, int l, uint32_t pix, uint32_t samp, int layer,
                        V3 p, V3 nu, V3 &q, V3 &w, float &cos_i, float &gl) {
  float u[4];
  rng4(P, pix, samp, (uint32_t)layer, 16u + (uint32_t)l, u);
  const V3 cor = v3(L.corner[0], L.corner[1], L.corner[2]);
  const V3 eu = v3(L.edge_u[0], L.edge_u[1], L.edge_u[2]), ev = v3(L.edge_v[0], L.edge_v[1], L.edge_v[2]);
  q = (cor + eu * u[0]) + ev * u[1];
  const V3 delta = p - q;
  const float d2 = dot(delta, delta);
  if (!(d2 > 0.0f)) return false;
  w = delta * (1.0f / __builtin_sqrtf(d2));
  const V3 nl = cross(eu, ev);
  const float area = __builtin_sqrtf(dot(nl, nl));
  const V3 nlu = nl * (1.0f / area);
  const float cos_l = dot(nlu, w);
  if (!(cos_l > 0.0f)) return false;
  cos_i = -dot(nu, w);
  if (!(cos_i > 0.0f)) return false;
  gl = (cos_l * area) / d2;
  return true;
}

// BlockBasedFrameBuffer<8,16>::blockPixelLocation (BlockBasedFrameBuffer.impl.cuh:125-143);
// K = frame_k of the request.
DR_HD bool block_pixel(const FrameK &K, int width, int height, int block, int pixel, int &x, int &y) {
  const int by = (int)udiv((uint32_t)block, K.nbw), bx = block - by * (int)K.nbw.d;
  const int oy = pixel / 8, ox = pixel - oy * 8;
  x = 8 * bx + ox;
  y = 16 * by + oy;
  return x < width && y < height;
}

}  // namespace dr
