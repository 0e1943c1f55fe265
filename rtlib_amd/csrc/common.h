// Shared host/device math for the rtpg path-tracing core (gfx950).
//
// Everything here is plain value arithmetic usable from host C++ (BVH build,
// SD-tree host logic) and from HIP kernels.  The random-number, canonical
// direction and ONB conventions follow RTLib's device math so that guide
// records land in the same quadtree cells as in the reference:
//   - float from 32 random bits: RTLib/Src/Ext/CUDA/Include/RTLib/Ext/CUDA/Math/Math.h:682-695
//   - ONB:                        .../Math/VectorFunction.h:5666-5697
//   - canonical_to_dir/dir_to_canonical: .../Math/VectorFunction.h:5698-5720
//   - cosine-weighted direction:  .../Math/Random.h:322-336
//   - point in triangle:          .../Math/Random.h:290-296
//   - pcg1d / xxhash32:           .../Math/Hash.h:13-18, 50-66
// The random stream itself is PCG32 (O'Neill 2014, pcg32_srandom_r /
// pcg32_random_r), one stream per pixel.
#pragma once
#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define RTPG_HD __host__ __device__ __forceinline__
#else
#define RTPG_HD inline
#endif

namespace rtpg {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.31830988618379067154f;
constexpr float k2Pi = 6.28318530717958647692f;
constexpr float kInv2Pi = 0.15915494309189533577f;
constexpr float kInv4Pi = 0.07957747154594766788f;

struct V3 {
  float x, y, z;
};

RTPG_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
RTPG_HD V3 v3(float s) { return V3{s, s, s}; }
RTPG_HD V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
RTPG_HD V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
RTPG_HD V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
RTPG_HD V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
RTPG_HD V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
RTPG_HD V3 operator*(float s, V3 a) { return V3{a.x * s, a.y * s, a.z * s}; }
RTPG_HD V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
RTPG_HD V3 operator/(V3 a, V3 b) { return V3{a.x / b.x, a.y / b.y, a.z / b.z}; }
RTPG_HD V3& operator+=(V3& a, V3 b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
RTPG_HD V3& operator*=(V3& a, V3 b) { a.x *= b.x; a.y *= b.y; a.z *= b.z; return a; }
RTPG_HD V3& operator*=(V3& a, float s) { a.x *= s; a.y *= s; a.z *= s; return a; }
RTPG_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RTPG_HD V3 cross(V3 a, V3 b) {
  return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
RTPG_HD float length(V3 a) { return sqrtf(dot(a, a)); }
RTPG_HD V3 normalize(V3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
RTPG_HD float avg3(V3 a) { return (a.x + a.y + a.z) / 3.0f; }
RTPG_HD float max3(V3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
RTPG_HD V3 vmin(V3 a, V3 b) { return V3{fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
RTPG_HD V3 vmax(V3 a, V3 b) { return V3{fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
RTPG_HD float comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
RTPG_HD bool finite3(V3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }

// ---------------------------------------------------------------- random
// PCG32 (XSH-RR 64/32).  `inc` is derived from the stream id and never stored.
constexpr uint64_t kPcgMult = 6364136223846793005ULL;

struct Pcg32 {
  uint64_t state;
  uint64_t inc;
  RTPG_HD uint32_t next() {
    uint64_t old = state;
    state = old * kPcgMult + inc;
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((~rot + 1u) & 31u));
  }
};

RTPG_HD uint64_t pcg_stream_inc(uint64_t stream) { return (stream << 1u) | 1u; }

// pcg32_srandom_r(initstate, initseq)
RTPG_HD Pcg32 pcg32_seed(uint64_t initstate, uint64_t stream) {
  Pcg32 r;
  r.state = 0u;
  r.inc = pcg_stream_inc(stream);
  r.next();
  r.state += initstate;
  r.next();
  return r;
}

// splitmix64 finaliser: decorrelates (seed, sample index) into a PCG start state.
RTPG_HD uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

// Start state of the path for (pixel, sample) under a global seed: the stream is
// the pixel, the position in the stream comes from the seed and the sample index.
RTPG_HD Pcg32 path_rng(uint64_t seed, uint32_t pixel, uint32_t sample) {
  return pcg32_seed(mix64(seed ^ mix64((uint64_t)sample)), (uint64_t)pixel);
}

// Stream of the recording filters for the vertex at `depth` of path (pixel,
// sample): separate from the path's own stream, so switching the filters on
// leaves every path unchanged.
RTPG_HD Pcg32 filter_rng(uint64_t seed, uint32_t pixel, uint32_t sample, uint32_t depth) {
  return pcg32_seed(mix64(seed ^ 0xD6E8FEB86659FD93ULL ^ mix64(((uint64_t)sample << 8) | depth)), (uint64_t)pixel);
}

// 23 random mantissa bits -> [0,1): the reference's to_uniform_float.
RTPG_HD float to_uniform_float(uint32_t v) {
  uint32_t bits = (v >> 9) | 0x3F800000u;
  float f;
#if defined(__HIP_DEVICE_COMPILE__)
  f = __uint_as_float(bits);
#else
  __builtin_memcpy(&f, &bits, 4);
#endif
  return f - 1.0f;
}

RTPG_HD float rnd(Pcg32& r) { return to_uniform_float(r.next()); }

RTPG_HD uint32_t fbits(float f) {
  uint32_t u;
#if defined(__HIP_DEVICE_COMPILE__)
  u = __float_as_uint(f);
#else
  __builtin_memcpy(&u, &f, 4);
#endif
  return u;
}
RTPG_HD float bitsf(uint32_t u) {
  float f;
#if defined(__HIP_DEVICE_COMPILE__)
  f = __uint_as_float(u);
#else
  __builtin_memcpy(&f, &u, 4);
#endif
  return f;
}

// ---------------------------------------------------------------- hashes
RTPG_HD uint32_t pcg1d(uint32_t input) {
  uint32_t state = input * 747796405u + 2891336453u;
  uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}

// 30-bit 3-D Morton code of 10-bit cell coordinates (x in bit 0, y bit 1, z bit 2).
RTPG_HD uint32_t part1by2_10(uint32_t x) {
  x &= 0x3FFu;
  x = (x | (x << 16)) & 0x030000FFu;
  x = (x | (x << 8)) & 0x0300F00Fu;
  x = (x | (x << 4)) & 0x030C30C3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}
RTPG_HD uint32_t compact1by2_10(uint32_t x) {
  x &= 0x09249249u;
  x = (x | (x >> 2)) & 0x030C30C3u;
  x = (x | (x >> 4)) & 0x0300F00Fu;
  x = (x | (x >> 8)) & 0x030000FFu;
  x = (x | (x >> 16)) & 0x000003FFu;
  return x;
}
RTPG_HD uint32_t morton3_30(uint32_t x, uint32_t y, uint32_t z) {
  return part1by2_10(x) | (part1by2_10(y) << 1) | (part1by2_10(z) << 2);
}

// 2-D Morton code of 16-bit direction-bin coordinates: x in even bits, y odd.
// Same interleave as the reference's Morton2Utils (OPX7UtilsMorton.h:160-225).
RTPG_HD uint32_t part1by1_16(uint32_t x) {
  x &= 0x0000FFFFu;
  x = (x | (x << 8)) & 0x00FF00FFu;
  x = (x | (x << 4)) & 0x0F0F0F0Fu;
  x = (x | (x << 2)) & 0x33333333u;
  x = (x | (x << 1)) & 0x55555555u;
  return x;
}
RTPG_HD uint32_t compact1by1_16(uint32_t x) {
  x &= 0x55555555u;
  x = (x | (x >> 1)) & 0x33333333u;
  x = (x | (x >> 2)) & 0x0F0F0F0Fu;
  x = (x | (x >> 4)) & 0x00FF00FFu;
  x = (x | (x >> 8)) & 0x0000FFFFu;
  return x;
}
RTPG_HD uint32_t morton2(uint32_t x, uint32_t y) { return (part1by1_16(y) << 1) + part1by1_16(x); }

// ---------------------------------------------------------------- frames
// Orthonormal basis around n; construction of VectorFunction.h:5666-5697.
struct Onb {
  V3 t, b, n;
};
RTPG_HD Onb make_onb(V3 normal) {
  Onb o;
  o.n = normalize(normal);
  if (fabsf(o.n.x) > fabsf(o.n.z)) {
    o.b = V3{-o.n.y, o.n.x, 0.0f};
  } else {
    o.b = V3{0.0f, -o.n.z, o.n.y};
  }
  o.b = normalize(o.b);
  o.t = normalize(cross(o.n, o.b));
  return o;
}
RTPG_HD V3 onb_local(const Onb& o, V3 d) { return d.x * o.t + d.y * o.b + d.z * o.n; }

RTPG_HD V3 canonical_to_dir(float px, float py) {
  float cos_t = 2.0f * px - 1.0f;
  float sin_t = sqrtf(fmaxf(1.0f - cos_t * cos_t, 0.0f));
  float phi = k2Pi * py;
  return V3{sin_t * cosf(phi), sin_t * sinf(phi), cos_t};
}
RTPG_HD void dir_to_canonical(V3 d, float& px, float& py) {
  if (!finite3(d)) {
    px = 0.0f;
    py = 0.0f;
    return;
  }
  float z = fminf(fmaxf(d.z, -1.0f), 1.0f);
  float phi = atan2f(d.y, d.x);
  while (phi < 0.0f) phi += k2Pi;
  px = (z + 1.0f) / 2.0f;
  py = phi * kInv2Pi;
}

RTPG_HD V3 cosine_dir_local(float u1, float u2) {
  float r = sqrtf(u1);
  float z = sqrtf(1.0f - u1);
  float t = 2.0f * u2 * kPi;
  return V3{r * cosf(t), r * sinf(t), z};
}

// ---------------------------------------------------------------- materials
enum MaterialType : uint32_t {
  kMatLambert = 0,  // albedo/pi
  kMatGGX = 1,      // GGX microfacet reflector, Schlick Fresnel with F0 = albedo
  kMatPhong = 2,    // reference "PHONG": Lambert + normalised Phong lobe
  kMatGlass = 3,    // reference dielectric with Schlick Fresnel (delta)
};

RTPG_HD float ggx_d(float cos_h, float a2) {
  float c2 = cos_h * cos_h;
  float den = c2 * (a2 - 1.0f) + 1.0f;
  return a2 / (kPi * den * den);
}
// Smith G1 for GGX.
RTPG_HD float ggx_g1(float cos_v, float a2) {
  float c = fabsf(cos_v);
  float c2 = c * c;
  return 2.0f * c / (c + sqrtf(a2 + (1.0f - a2) * c2));
}
RTPG_HD V3 schlick3(V3 f0, float c) {
  float m = 1.0f - fminf(fmaxf(c, 0.0f), 1.0f);
  float m5 = m * m * m * m * m;
  return f0 + (v3(1.0f) - f0) * m5;
}

// BSDF value f(wo, wi) (without the cosine) and sampling pdf of wi, both
// for the non-delta materials.  n is the shading normal facing wo.
struct BsdfEval {
  V3 f;
  float pdf;
};

// Normalised Phong lobe pdf (SimpleTrace.cu:24-44).  For shininess 0 the
// reference's own powf wrapper returns (x == 0) instead of 1
// (CUDA/Math/Math.h:615-620); that quirk is kept.
RTPG_HD float phong_lobe_pdf(V3 dir, V3 refl, float shin) {
  float rc = fmaxf(dot(refl, dir), 0.0f);
  float p = (shin != 0.0f) ? powf(rc, shin) : (rc == 0.0f ? 1.0f : 0.0f);
  return (shin + 1.0f) * p * kInv2Pi;
}

RTPG_HD BsdfEval bsdf_eval(uint32_t type, V3 albedo, float param, V3 spec, V3 n, V3 wo, V3 wi) {
  BsdfEval e;
  float cos_i = dot(n, wi);
  if (type == kMatGGX) {
    float cos_o = dot(n, wo);
    if (cos_i <= 0.0f || cos_o <= 0.0f) {
      e.f = v3(0.0f);
      e.pdf = 0.0f;
      return e;
    }
    float a = fmaxf(param, 1e-3f);
    float a2 = a * a;
    V3 h = normalize(wi + wo);
    float cos_h = dot(n, h);
    float d = ggx_d(cos_h, a2);
    float g = ggx_g1(cos_i, a2) * ggx_g1(cos_o, a2);
    float oh = fmaxf(dot(wo, h), 1e-8f);
    V3 fr = schlick3(albedo, oh);
    e.f = fr * (d * g / (4.0f * cos_i * cos_o));
    e.pdf = d * cos_h / (4.0f * oh);
    return e;
  }
  if (type == kMatPhong) {
    V3 refl = normalize(wo * -1.0f + n * (2.0f * dot(n, wo)));
    float shin = param;
    float aver_diff = 1.0e-10f + avg3(albedo);
    float aver_spec = 1.0e-10f + avg3(spec);
    float sel = aver_diff / (aver_diff + aver_spec);
    float lobe = phong_lobe_pdf(wi, refl, shin);
    e.f = albedo * kInvPi + spec * ((shin + 2.0f) / (shin + 1.0f)) * lobe;
    e.pdf = sel * fmaxf(cos_i * kInvPi, 0.0f) + (1.0f - sel) * lobe;
    if (cos_i <= 0.0f) e.f = v3(0.0f);
    return e;
  }
  // Lambert
  if (cos_i <= 0.0f) {
    e.f = v3(0.0f);
    e.pdf = 0.0f;
    return e;
  }
  e.f = albedo * kInvPi;
  e.pdf = cos_i * kInvPi;
  return e;
}

// Draws wi for the non-delta materials; consumes exactly two (Lambert, GGX)
// or three (Phong) random numbers.
RTPG_HD V3 bsdf_sample(uint32_t type, V3 albedo, float param, V3 spec, V3 n, V3 wo, Pcg32& rng) {
  if (type == kMatGGX) {
    float u1 = rnd(rng), u2 = rnd(rng);
    float a = fmaxf(param, 1e-3f);
    float a2 = a * a;
    float cos_h = sqrtf((1.0f - u1) / (1.0f + (a2 - 1.0f) * u1));
    float sin_h = sqrtf(fmaxf(1.0f - cos_h * cos_h, 0.0f));
    float phi = k2Pi * u2;
    Onb o = make_onb(n);
    V3 h = onb_local(o, V3{sin_h * cosf(phi), sin_h * sinf(phi), cos_h});
    return h * (2.0f * dot(wo, h)) - wo;
  }
  if (type == kMatPhong) {
    float u_sel = rnd(rng), u1 = rnd(rng), u2 = rnd(rng);
    float aver_diff = 1.0e-10f + avg3(albedo);
    float aver_spec = 1.0e-10f + avg3(spec);
    float sel = aver_diff / (aver_diff + aver_spec);
    if (u_sel < sel) {
      return onb_local(make_onb(n), cosine_dir_local(u1, u2));
    }
    V3 refl = normalize(wo * -1.0f + n * (2.0f * dot(n, wo)));
    float shin = param;
    float cos_t = powf(u1, 1.0f / (shin + 1.0f));
    float sin_t = sqrtf(fmaxf(1.0f - cos_t * cos_t, 0.0f));
    float phi = k2Pi * u2;
    return onb_local(make_onb(refl), V3{sin_t * cosf(phi), sin_t * sinf(phi), cos_t});
  }
  float u1 = rnd(rng), u2 = rnd(rng);
  return onb_local(make_onb(n), cosine_dir_local(u1, u2));
}

RTPG_HD float power_heuristic(float a, float b) {
  float a2 = a * a;
  float b2 = b * b;
  return (a2 + b2) > 0.0f ? a2 / (a2 + b2) : 0.0f;
}

}  // namespace rtpg
