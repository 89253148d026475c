// The reference's own render plugins on the device (DR_PLUGINS_REFERENCE): shaders
// DumbBasedShader, SimpleStochasticShader, FresnelShader and DefaultShaderGeneric with
// their textures (ColoredTexture colour, diffuse, normal-map and alpha-cutout maps), lights
// SimpleDirectionalLight, SimpleSoftDirectionalLight, SphericalLightEmitter and Spotlight.  Each function keeps the
// reference's expression order, so a host evaluation of the same code gives the
// same bits.
//
// Random numbers are Philox uniforms of (pixel, sample, layer, purpose):
//   purpose 1: u0, u1 = DumbRand::pointOnSphere, u2 = Russian roulette,
//              u3 = DumbRand::getBool of DumbBasedShader;
//   purpose 16 + l: light l;  purpose 32 + i: try i of DumbBasedShader's specular lobe.
#pragma once
#include "dr_shade.h"

namespace dr {

constexpr float kPiRef = 3.14159265359f;   // PI, Vector3.h:8
constexpr int kSpecTries = 224;            // tries of the specular lobe's rejection loop
constexpr int kMaxTransmit = 16;           // surfaces a light photon may pass through
constexpr int kFaceWords = 24;             // per triangle: a, b, c, the three vertex normals, then u, v of a, b, c

DR_HD bool is_reference_kind(int kind) { return kind >= DR_DUMB_BASED; }
DR_HD V3 neg(V3 a) { return v3(-a.x, -a.y, -a.z); }

// ---------------------------------------------------------------- pow
// x^y for the shaders' pow calls.  log2 of the mantissa m in [sqrt(1/2), sqrt(2)) by the
// series ln m = 2 atanh t = 2 (t + t^3/3 + t^5/5 + ...), t = (m-1)/(m+1), |t| <= 0.172;
// 2^f for f in [0,1) by the Taylor series of e^(f ln 2) to the 9th power; 2^k exactly by
// ldexp.  Negative bases follow C: integer y gives the signed result, other y NaN.
DR_HD float log2_pos(float x) {
  int e;
  float m = __builtin_frexpf(x, &e);
  if (m < 0.707106781f) { m = m * 2.0f; e = e - 1; }
  const float t = (m - 1.0f) / (m + 1.0f);
  const float t2 = t * t;
  const float s = t * (2.0f + t2 * (0.666666667f + t2 * (0.4f + t2 * (0.285714286f + t2 * (0.222222222f +
                  t2 * 0.181818182f)))));
  return (float)e + s * 1.44269504f;
}
DR_HD float exp2_ref(float z) {
  if (z != z) return z;
  if (z >= 128.0f) return __builtin_inff();
  if (z < -150.0f) return 0.0f;
  const float k = __builtin_floorf(z);
  const float a = (z - k) * 0.693147181f;
  const float p = 1.0f + a * (1.0f + a * (0.5f + a * (0.166666667f + a * (0.0416666667f + a * (0.00833333333f +
                  a * (0.00138888889f + a * (0.000198412698f + a * (2.48015873e-05f + a * 2.75573192e-06f))))))));
  return __builtin_ldexpf(p, (int)k);
}
DR_HD float pow_ref(float x, float y) {
  if (y == 0.0f || x == 1.0f) return 1.0f;
  if (x != x || y != y) return x + y;
  bool neg = false;
  if (x < 0.0f) {
    if (__builtin_floorf(y) != y) return __builtin_nanf("");
    neg = (__builtin_fabsf(y) < 16777216.0f) && ((((long long)y) & 1) != 0);
    x = -x;
  }
  float r;
  if (x == 0.0f) r = (y < 0.0f) ? __builtin_inff() : 0.0f;
  else if (x == __builtin_inff()) r = (y < 0.0f) ? 0.0f : __builtin_inff();
  else r = exp2_ref(y * log2_pos(x));
  return neg ? -r : r;
}

// DumbRand::pointOnSphere (DumbRand.impl.cuh:70-77) with theta = 2 pi ua, cos phi = 1 - 2 ub
DR_HD V3 point_on_sphere(float ua, float ub, float radius) {
  float s, c;
  sincos_2pi(ua, s, c);
  const float cosPhi = (1.0f - (2.0f * ub));
  const float sinPhi = __builtin_sqrtf(1.0f - (cosPhi * cosPhi));
  return v3(((sinPhi * c) * radius), ((sinPhi * s) * radius), (cosPhi * radius));
}

// Vector3::reflection (Vector3.impl.h:418-421)
DR_HD V3 reflection(V3 v, V3 n) {
  const float mul = ((2.0f * dot(v, n)) / dot(n, n));
  return v3(v.x - (n.x * mul), v.y - (n.y * mul), v.z - (n.z * mul));
}
// Vector3::angleCos (Vector3.impl.h:292-294, AngleCosWith :29)
DR_HD float angle_cos(V3 u, V3 v) { return dot(u, v) / __builtin_sqrtf(dot(u, u) * dot(v, v)); }
// DumbTools::fresnel (DumbStructs.cuh:136-140)
DR_HD void fresnel(const float F[3], V3 n, V3 ray, float out[3]) {
  const float val = (1.0f - dot(n, ray));
  const float sq = (val * val);
  const float s = (sq * sq) * val;
  for (int c = 0; c < 3; c++) out[c] = F[c] + ((1.0f - F[c]) * s);
}
// Vector3::normalOn (Vector3.impl.h:364-366, NormalOn :33)
DR_HD V3 normal_on(V3 t, V3 v) { return v * (dot(t, v) / dot(v, v)); }
// Triangle::getMasses (Triangle.impl.h:248-277)
DR_HD V3 get_masses(V3 a, V3 b, V3 c, V3 center) {
  const V3 ab = b - a;
  V3 bc = c - b;
  V3 ae = ab - normal_on(ab, bc);
  const V3 ax = center - a;
  V3 ad = normal_on(ax, ae);
  if (ae.x < 0) { ad.x = -ad.x; ae.x = -ae.x; }
  if (ae.y < 0) { ad.y = -ad.y; ae.y = -ae.y; }
  if (ae.z < 0) { ad.z = -ad.z; ae.z = -ae.z; }
  float div = ad.x + ad.y + ad.z;
  if (div == 0) return v3(1, 0, 0);
  const float g = (ae.x + ae.y + ae.z) / div;
  float t;
  V3 by = (a + ax * g) - b;
  if (bc.x < 0) { bc.x = -bc.x; by.x = -by.x; }
  if (bc.y < 0) { bc.y = -bc.y; by.y = -by.y; }
  if (bc.z < 0) { bc.z = -bc.z; by.z = -by.z; }
  div = bc.x + bc.y + bc.z;
  if (div == 0) t = 0;
  else t = (by.x + by.y + by.z) / div;
  const float cc = t, bb = (1 - t), aa = (g - 1);
  return divf(v3(aa, bb, cc), aa + bb + cc);
}

// ---------------------------------------------------------------- textures
// Textures of the reference shaders: the tables of dr_render_params.
struct TexCtx {
  const dr_material_maps *maps;
  const dr_texture *tex;
  const float4 *texels;
};
DR_HD TexCtx tex_ctx(const dr_render_params &P) { return TexCtx{P.maps, P.textures, (const float4 *)P.texels}; }
// The material's maps, or nullptr (dr_material.reserved[0] = 1 + index).
DR_HD const dr_material_maps *material_maps(const TexCtx &T, const dr_material &M) {
  return (M.reserved[0] > 0 && T.maps) ? &T.maps[M.reserved[0] - 1] : nullptr;
}

// Texture::operator()(Vector2) (Texture.impl.cuh:78-90): wrap to [0,1) through the integer
// casts, flip y, then the nearest texel of the 2x2 cell (FILTER_NONE) or bilinear weights.
DR_HD float4 texture_sample(const dr_texture &T, const float4 *__restrict__ texels, float px, float py) {
  float x = (px + ((float)(((int)(-px)) + 2)));
  x -= ((float)((int)x));
  float y = (py + ((float)(((int)(-py)) + 2)));
  y -= ((float)((int)y));
  y = (1.0f - y);
  const uint32_t w = (uint32_t)T.width, h = (uint32_t)T.height;
  const float posX = (x * ((float)w)), posY = (y * ((float)h));
  const uint32_t minX = (((uint32_t)posX) % w), minY = (((uint32_t)posY) % h);
  const uint32_t maxX = ((minX + 1) % w), maxY = ((minY + 1) % h);
  const float offX = (posX - ((float)minX)), offY = (posY - ((float)minY));
  const float4 *D = texels + T.offset;
  if (T.filter == 0) return D[((offY <= 0.5f) ? minY : maxY) * w + ((offX <= 0.5f) ? minX : maxX)];
  const float4 c00 = D[minY * w + minX], c01 = D[maxY * w + minX], c10 = D[minY * w + maxX], c11 = D[maxY * w + maxX];
  const float w00 = ((1.0f - offX) * (1.0f - offY)), w01 = ((1.0f - offX) * offY);
  const float w10 = (offX * (1.0f - offY)), w11 = (offX * offY);
  float4 r;
  r.x = (((c00.x * w00) + (c01.x * w01)) + (c10.x * w10)) + (c11.x * w11);
  r.y = (((c00.y * w00) + (c01.y * w01)) + (c10.y * w10)) + (c11.y * w11);
  r.z = (((c00.z * w00) + (c01.z * w01)) + (c10.z * w10)) + (c11.z * w11);
  r.w = (((c00.w * w00) + (c01.w * w01)) + (c10.w * w10)) + (c11.w * w11);
  return r;
}
// The texture sample of a map at texture coordinate (u, v): coord = (uv ^ tiling) + offset
// (ColoredTexture::operator(), DumbStructs.cuh:87-91).
DR_HD float4 map_sample(const TexCtx &T, const dr_texture_map &m, float u, float v) {
  return texture_sample(T.tex[m.texture], T.texels, (u * m.tiling[0]) + m.offset[0], (v * m.tiling[1]) + m.offset[1]);
}
// A ColoredTexture's colour at (u, v): the colour, times the texture sample when it has one.
DR_HD void map_color(const TexCtx &T, const dr_texture_map *m, const float col[3], float u, float v, float out[3]) {
  if (m == nullptr || m->texture < 0) {
    for (int c = 0; c < 3; c++) out[c] = col[c];
    return;
  }
  const float4 t = map_sample(T, *m, u, v);
  out[0] = col[0] * t.x;
  out[1] = col[1] * t.y;
  out[2] = col[2] * t.z;
}

// A surface point as the reference shaders see it.
struct RefHit {
  V3 p;
  V3 ng;   // Triangle::normal = (c - a) x (b - a) (Triangle.impl.h:233-235)
  V3 nI;   // BakedTriFace::norm.massCenter(getMasses(p)) (Triangle.impl.h:318-320)
  V3 ns;   // ColoredTexture::getNormal of normalColor (DumbStructs.cuh:107-131): nI normalised,
           // or turned by the normal map
  float tu, tv;   // BakedTriFace::tex.massCenter(getMasses(p))
};
__device__ __forceinline__ RefHit ref_hit(const float *__restrict__ faces, int h, V3 p, const TexCtx &T,
                                          const dr_material_maps *mp) {
  const float *F = faces + (size_t)kFaceWords * h;
  const V3 a = v3(F[0], F[1], F[2]), b = v3(F[3], F[4], F[5]), c = v3(F[6], F[7], F[8]);
  RefHit H;
  H.p = p;
  H.ng = cross(c - a, b - a);
  const V3 m = get_masses(a, b, c, p);
  const V3 na = v3(F[9], F[10], F[11]), nb = v3(F[12], F[13], F[14]), nc = v3(F[15], F[16], F[17]);
  H.nI = (na * m.x + nb * m.y) + nc * m.z;
  H.ns = normalized(H.nI);
  H.tu = (F[18] * m.x + F[20] * m.y) + F[22] * m.z;
  H.tv = (F[19] * m.x + F[21] * m.y) + F[23] * m.z;
  if (mp != nullptr && mp->map[2].texture >= 0) {
    // tangent from the texture coordinates (DumbStructs.cuh:111-127)
    float du, dv;
    V3 vd;
    const V3 pm = (a * m.x + b * m.y) + c * m.z;
    if (m.x < 0.999f) { du = F[18] - H.tu; dv = F[19] - H.tv; vd = a - pm; }
    else { du = F[20] - H.tu; dv = F[21] - H.tv; vd = b - pm; }
    const V3 n = H.ns;
    const V3 vco = normalized(cross(vd, n));
    vd = cross(vco, n);
    const V3 u = normalized((vd * dv) - (vco * du));
    const V3 v = cross(u, n);
    const float4 rgb = map_sample(T, mp->map[2], H.tu, H.tv);
    const float mx = (mp->normal_color[1] * rgb.z);
    const float my = (mp->normal_color[0] * ((rgb.x * 2.0f) - 1.0f));
    const float mz = (mp->normal_color[2] * ((rgb.y * 2.0f) - 1.0f));
    H.ns = normalized(((n * mx) + (u * my)) + (v * mz));
  }
  return H;
}
// DumbBasedShader's alpha cutout (DumbBasedShader.impl.cuh:29, :68): the cutout texture's
// alpha at the hit is at most the threshold.
DR_HD bool alpha_cut(const TexCtx &T, const dr_material_maps *mp, const RefHit &H) {
  if (mp == nullptr || mp->map[3].texture < 0) return false;
  return mp->alpha_threshold >= map_sample(T, mp->map[3], H.tu, H.tv).w;
}

// getReflectedColor of a reference shader for a photon (direction pd, colour pc) toward
// the observer direction ob; direct = PHOTON_TYPE_DIRECT_ILLUMINATION.
DR_HD void reflect_ref(const dr_material &M, const RefHit &H, V3 pd, const float pc[3], V3 ob, bool direct, int stype,
                       const TexCtx &T, float out[3]) {
  const dr_material_maps *mp = material_maps(T, M);
  if (M.kind == DR_DUMB_BASED) {                      // DumbBasedShader.impl.cuh:63-109
    if (alpha_cut(T, mp, H)) {
      const bool through = dot(normalized(ob), normalized(pd)) >= 0.9999f;
      for (int c = 0; c < 3; c++) out[c] = through ? pc[c] : 0.0f;
      return;
    }
    if (dot(normalized(H.ng), pd) > 0.0f && !direct) {
      for (int c = 0; c < 3; c++) out[c] = pc[c];
      return;
    }
    if (dot(H.nI, ob) < 0.0f) { out[0] = out[1] = out[2] = 0.0f; return; }
    const float spec = M.p[0], metal = M.p[1];
    const V3 n = H.ns;
    const V3 wi = neg(normalized(pd));
    const bool brdf_m = direct || stype == 1;
    float brdf[3] = {0.0f, 0.0f, 0.0f};
    if (brdf_m) {
      const V3 w0 = normalized(ob);
      const V3 wh = normalized(w0 + wi);
      float fcol[3], fwi[3];
      map_color(T, mp ? &mp->map[0] : nullptr, M.albedo, H.tu, H.tv, fcol);
      fresnel(fcol, wh, wi, fwi);
      float mult = 1.0f;
      if (direct) {
        const float nwh = dot(n, wh);
        const float dwh = ((spec + 2.0f) / (2.0f * kPiRef)) * pow_ref(nwh, spec);
        const float nw0 = dot(n, w0), nwi = dot(n, wi), w0wh = dot(w0, wh);
        const float g = smin(1.0f, smin(((2.0f * nwh) * nw0) / w0wh, ((2.0f * nwh) * nwi) / w0wh));
        mult = (dwh * g) / ((4.0f * nw0) * nwi);
      }
      for (int c = 0; c < 3; c++) brdf[c] = smax(0.0f, smin(fwi[c] * mult, 1.0f));
    }
    const bool diff_m = direct || stype == 2;
    float dif[3] = {0.0f, 0.0f, 0.0f};
    if (diff_m) {
      float dcol[3];
      map_color(T, mp ? &mp->map[1] : nullptr, M.diffuse, H.tu, H.tv, dcol);
      const float nwi = dot(n, wi);
      for (int c = 0; c < 3; c++) dif[c] = dcol[c] * nwi;
    }
    float col[3];
    if (brdf_m && diff_m) {
      for (int c = 0; c < 3; c++) col[c] = smax(0.0f, smin((brdf[c] * metal) + dif[c], 1.0f));
    } else if (brdf_m) {
      for (int c = 0; c < 3; c++) col[c] = brdf[c];
    } else {
      const float inv = (1.0f / (1.0f - metal));
      for (int c = 0; c < 3; c++) col[c] = dif[c] * inv;
    }
    for (int c = 0; c < 3; c++) out[c] = col[c] * pc[c];
  } else if (M.kind == DR_SIMPLE_STOCHASTIC) {        // SimpleStochasticShader.impl.cuh:25-34
    const V3 refl = reflection(pd, H.nI);
    float rc = angle_cos(refl, H.nI);
    if (rc < 0.0f) rc = 0.0f;
    float oc = angle_cos(ob, refl);
    if (oc < 0.0f) oc = 0.0f;
    const float s = (M.p[0] * rc) + (M.p[1] * pow_ref(oc, M.p[2]));
    for (int c = 0; c < 3; c++) out[c] = (pc[c] * M.albedo[c]) * s;
  } else if (M.kind == DR_FRESNEL) {                  // FresnelShader.impl.cuh:14-19
    const V3 light = neg(normalized(pd));
    float fcol[3], f[3];
    map_color(T, mp ? &mp->map[0] : nullptr, M.albedo, H.tu, H.tv, fcol);
    fresnel(fcol, H.nI, light, f);
    for (int c = 0; c < 3; c++) out[c] = f[c] * pc[c];
  } else {                                            // DefaultShaderGeneric::cast, DefaultShader.impl.cuh:15-33
    const V3 cam = (H.p + ob) - H.p;
    const V3 refl = reflection(pd, H.nI);
    const float cl = smax(angle_cos(refl, H.nI), 0.0f);
    const float oc = smax(angle_cos(refl, cam), 0.0f);
    const float s = (M.p[0] * cl) + (M.p[1] * pow_ref(oc, M.p[2]));
    for (int c = 0; c < 3; c++) out[c] = (pc[c] * M.albedo[c]) * s;
  }
}

// requestIndirectSamples of a reference shader (uniforms u of purpose 1).
DR_HD bool sample_ref(const dr_render_params &P, const dr_material &M, const RefHit &H, V3 d, uint32_t pix,
                      uint32_t samp, int layer, const float u[4], V3 &dir, int &type) {
  if (M.kind == DR_DUMB_BASED) {                      // DumbBasedShader.impl.cuh:18-61
    const TexCtx T = tex_ctx(P);
    if (alpha_cut(T, material_maps(T, M), H)) { dir = d; type = 3; return true; }   // cut out: straight through
    V3 normal = normalized(H.ng);
    if (u[3] <= M.p[1]) {
      const V3 n = H.ns;
      if (dot(n, d) >= 0.0f) return false;
      const V3 r = normalized(reflection(d, n));
      if (dot(r, normal) < 0.0f) normal = normal * -1.0f;
      V3 a = cross(r, v3(1.0f, 0.0f, 0.0f));
      if (dot(normal, normal) < kVecEps) a = cross(r, v3(0.0f, 1.0f, 0.0f));
      a = normalized(a);
      const V3 b = cross(r, a);
      const float inv_spec = (1.0f / M.p[0]);
      for (int i = 0; i < kSpecTries; i++) {
        float v[4];
        rng4(P, pix, samp, (uint32_t)layer, 32u + (uint32_t)i, v);
        const float cosine = pow_ref(v[0], inv_spec);
        const float sinus = __builtin_sqrtf(1.0f - (cosine * cosine));
        float as, ac;
        sincos_2pi(v[1], as, ac);
        const float angleSin = __builtin_sqrtf(1.0f - (ac * ac));
        const V3 sd = (r * cosine) + ((a * ac) + (b * angleSin)) * sinus;
        if (dot(sd, normal) >= 0.0f) { dir = sd; type = 1; return true; }
      }
      return false;
    }
    V3 sd = point_on_sphere(u[0], u[1], 1.0f);
    if (dot(sd, normal) < 0.0f) sd = sd * -1.0f;
    dir = sd;
    type = 2;
    return true;
  }
  if (M.kind == DR_SIMPLE_STOCHASTIC) {               // SimpleStochasticShader.impl.cuh:13-23
    V3 sd = point_on_sphere(u[0], u[1], 1.0f);
    if (dot(sd, H.ng) < 0.0f) sd = neg(sd);
    const V3 refl = reflection(d, H.nI);
    dir = normalized((sd * M.p[0]) + (refl * M.p[1]));
    type = 0;
    return true;
  }
  if (M.kind == DR_DEFAULT_SHADER) {                  // DefaultShader.impl.cuh:35-43
    dir = reflection(d, H.nI);
    type = 0;
    return true;
  }
  return false;                                       // FresnelShader.impl.cuh:11-13: no samples
}

// Lights::Spotlight flags (Spotlight.cuh:28-32), kept in dr_light.reserved.
constexpr int kSpotSampleMask = 31, kSpotCastShadows = 32, kSpotEmitFromSphere = 64;

// Photons a light sends per vertex (Spotlight: its sample count, at most 16; others 1).
DR_HD int light_samples(const dr_light &L) { return L.kind == DR_SPOT_LIGHT ? (L.reserved & kSpotSampleMask) : 1; }
// Whether the light's photons are tested for shadows (Light::getVertexPhotons castShadows).
DR_HD bool light_casts_shadows(const dr_light &L) { return L.kind != DR_SPOT_LIGHT || (L.reserved & kSpotCastShadows); }

// getVertexPhotons of light L (index l), photon k of its samples, for p: photon ray
// (q, w), colour pc, and whether w is unit length.  false: no photon.  Photon k of a
// multi-sample light draws its uniforms with k in bits 16-19 of the Philox counter.
DR_HD bool photon_ref(const dr_render_params &P, const dr_light &L, int l, int k, uint32_t pix, uint32_t samp,
                      int layer, V3 p, V3 &q, V3 &w, float pc[3], bool &unit) {
  float u[4];
  rng4(P, pix, samp, (uint32_t)layer | ((uint32_t)k << 8), 16u + (uint32_t)l, u);
  if (L.kind == DR_SPOT_LIGHT) {                      // Lights::Spotlight, Spotlight.impl.cuh:41-90
    const V3 dir = v3(L.edge_u[0], L.edge_u[1], L.edge_u[2]);
    const V3 pos0 = v3(L.corner[0], L.corner[1], L.corner[2]);
    const float size = L.radius, inner = L.edge_v[0], outer = L.edge_v[1], falloff = L.edge_v[2];
    const bool sphere = (L.reserved & kSpotEmitFromSphere) != 0;
    const V3 nondir = (dir.z > 0.9f) ? v3(1.0f, 0.0f, 0.0f) : v3(0.0f, 0.0f, 1.0f);
    const V3 ver = normalized(cross(dir, nondir));
    const V3 hor = cross(dir, ver);
    V3 pos;
    if (sphere) {
      pos = point_on_sphere(u[0], u[1], size);
      pos = pos * u[2];
      pos = pos + pos0;
    } else {
      float sa, ca;
      sincos_2pi(u[0], sa, ca);                       // angle = range(0, 2*PI)
      pos = pos0 + (((ver * ca) + (hor * sa)) * size) * __builtin_sqrtf(u[1]);
    }
    const V3 delta = p - pos;
    const float d2 = (delta.x * delta.x + delta.y * delta.y) + delta.z * delta.z;
    if (d2 < kVecEps) return false;
    const float dist = __builtin_sqrtf(d2);
    const float base = size * size;
    const float spread = __builtin_fmaxf(1.0f - outer, kVecEps);
    const float surface = (sphere ? 0.0f : base) + (d2 * spread);
    const float scale = base / (((float)(L.reserved & kSpotSampleMask)) * surface);
    for (int c = 0; c < 3; c++) pc[c] = L.emission[c] * scale;
    const V3 wd = delta * (1.0f / dist);
    const float cosine = dot(wd, dir);
    if (cosine < inner) {
      if (cosine < outer) return false;
      const float a = ((inner - cosine) / (inner - outer));
      const float b = (1.0f - a);
      const float f = pow_ref(((a * b) * b) + (b * (1.0f - (a * a))), falloff);
      for (int c = 0; c < 3; c++) pc[c] *= f;
    }
    q = pos;
    w = wd;
    unit = true;
    return true;
  }
  if (L.kind == DR_DIRECTIONAL_LIGHT) {               // SimpleDirectionalLight.impl.cuh:13-16
    const V3 dir = v3(L.edge_u[0], L.edge_u[1], L.edge_u[2]);
    q = p - dir * L.dist;
    w = dir;
    for (int c = 0; c < 3; c++) pc[c] = L.emission[c];
    unit = true;
    return true;
  }
  if (L.kind == DR_SOFT_DIRECTIONAL_LIGHT) {          // SimpleSoftDirectionalLight.impl.cuh:12-19
    const V3 dir = v3(L.edge_u[0], L.edge_u[1], L.edge_u[2]);
    const V3 off = point_on_sphere(u[0], u[1], L.radius * u[2]);
    const V3 origin = (p - dir * L.dist) + off;
    q = origin;
    w = p - origin;
    for (int c = 0; c < 3; c++) pc[c] = L.emission[c];
    unit = false;
    return true;
  }
  if (L.kind == DR_SPHERICAL_LIGHT) {                 // SphericalLightEmitter.impl.cuh:10-18
    V3 e = point_on_sphere(u[0], u[1], L.radius);
    e = e * u[2];
    e = e + v3(L.corner[0], L.corner[1], L.corner[2]);
    const V3 delta = p - e;
    q = e;
    w = normalized(delta);
    const float inv = (1.0f / dot(delta, delta));
    for (int c = 0; c < 3; c++) pc[c] = L.emission[c] * inv;
    unit = true;
    return true;
  }
  // area quad: colour E * cos_l * A / d^2
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
  const float gl = (cos_l * area) / d2;
  for (int c = 0; c < 3; c++) pc[c] = L.emission[c] * gl;
  unit = true;
  return true;
}

}  // namespace dr
