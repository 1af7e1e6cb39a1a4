// th_integrator.h -- the path estimator of the reference GPU kernel on gfx950.
//
// Restates shaders/pathtrace_spheres.wgsl:390-547 and 805-937 operation by operation
// (draw order, constants, evaluation order), on top of the traversal in th_device.h.
// Sampling accelerators keep the result identical to the reference's linear scans:
// emitter selection is a binary search for the first cumulative weight >= target
// (the scan of wgsl:394-400 stops at the same index), and the environment CDF search
// is the same lower-bound search as wgsl:352-361.
#pragma once
#include "th_wide.h"

namespace thd {

template <bool COUNT>
struct LaneCounters {
  uint32_t paths = 0, closest_rays = 0, shadow_rays = 0, nodes = 0, primrefs = 0, tris = 0,
           spheres = 0, closest_hits = 0, tri_hits = 0, lights_scanned = 0, env_lookups = 0,
           env_samples = 0, c_nodes = 0, c_primrefs = 0, c_tris = 0, c_spheres = 0;
  __device__ __forceinline__ void add_ray(const RayStats<COUNT>& r, bool shadow) {
    if (!COUNT) return;
    nodes += r.nodes;
    primrefs += r.primrefs;
    tris += r.tris;
    spheres += r.spheres;
    if (shadow) {
      shadow_rays++;
    } else {  // closest-hit rays alone as well
      closest_rays++;
      c_nodes += r.nodes;
      c_primrefs += r.primrefs;
      c_tris += r.tris;
      c_spheres += r.spheres;
    }
  }
  __device__ __forceinline__ void flush(unsigned long long* g) {
    if (!COUNT) return;
    uint32_t v[16] = {paths, closest_rays, shadow_rays, nodes, primrefs, tris, spheres,
                      closest_hits, tri_hits, lights_scanned, env_lookups, env_samples,
                      c_nodes, c_primrefs, c_tris, c_spheres};
#pragma unroll
    for (int i = 0; i < 16; ++i)
      if (v[i]) atomicAdd(g + i, (unsigned long long)v[i]);
  }
};

struct EnvSample {
  v3 dir, rad;
  float pdf;
};

template <int TMODE, bool COUNT>
struct Integrator {
  const DevScene& S;
  const TH_GpuRenderParams& P;
  const float* srgb_lut;
  Stack& st;
  LaneCounters<COUNT>& lc;
  const Child4* lds_nodes = nullptr;  // LDS copy of the top 4-wide nodes (quad traversal)
  uint32_t cached = 0;

  // closest hit; false = no surface within MAX_DISTANCE (wgsl:817-819)
  __device__ bool trace(const v3& o, const v3& d, Hit* h) {
    RayStats<COUNT> rs;
    float t;
    uint32_t slot;
    bool linear = S.node_count == 0u;
    if (linear) {
      t = closest_linear<COUNT>(S, P.sphere_count, P.triangle_count, o, d, TH_FLT_MAX, rs, &slot);
    } else if (TMODE == 0) {
      traverse<0, COUNT>(S, o, d, TH_FLT_MAX, st, rs, &t, &slot);
    } else if (TMODE == 1) {
      traverse<1, COUNT>(S, o, d, TH_FLT_MAX, st, rs, &t, &slot);
    } else {
      traverse_quad<1>(S, (const Child4*)S.nodes4, S.root4, o, d, TH_FLT_MAX, st, &t, &slot,
                       lds_nodes, cached);
    }
    lc.add_ray(rs, false);
    if (slot == THD_NO_SLOT || t > TH_MAX_DISTANCE) return false;
    if (linear)
      finalize_linear(S, slot, o, d, t, h);
    else
      finalize_slot(S, slot, o, d, t, h);
    if (COUNT) {
      lc.closest_hits++;
      bool tri = linear ? ((slot & THD_KIND_TRI) != 0u)
                        : ((fbits(S.slots[3u * (size_t)slot + 1u].w) & THD_KIND_TRI) != 0u);
      if (tri) lc.tri_hits++;
    }
    return true;
  }

  // true when the reference's closest hit along (o, d) lies at t < limit
  __device__ bool occluded(const v3& o, const v3& d, float limit) {
    RayStats<COUNT> rs;
    float t;
    uint32_t slot;
    if (S.node_count == 0u) {
      t = closest_linear<COUNT>(S, P.sphere_count, P.triangle_count, o, d, limit, rs, &slot);
    } else if (TMODE == 0) {
      traverse<0, COUNT>(S, o, d, TH_FLT_MAX, st, rs, &t, &slot);
    } else if (TMODE == 1) {
      traverse<2, COUNT>(S, o, d, limit, st, rs, &t, &slot);
    } else {
      traverse_quad<2>(S, (const Child4*)S.nodes4, S.root4, o, d, limit, st, &t, &slot, lds_nodes,
                       cached);
    }
    lc.add_ray(rs, true);
    return t < limit;
  }

  // One traversal step of an octet (two quads of one path, octet kernels): the quad with
  // shadow == false looks for the closest hit along (o, d) when `ext`; the other quad
  // tests (so, sdir) for an occluder before `slim` when `shd`.  Both run in the same
  // loop, so the two rays cost one traversal's latency.  Afterwards both quads hold the
  // closest hit (*t, *slot) and the occlusion result.
  __device__ bool octet_step(bool shadow, bool ext, const v3& o, const v3& d, bool shd, const v3& so,
                             const v3& sdir, float slim, float* t, uint32_t* slot) {
    float tt;
    uint32_t ss;
    bool r = traverse_quad_rt(S, (const Child4*)S.nodes4, S.root4, shadow ? shd : ext, shadow,
                              shadow ? so : o, shadow ? sdir : d, shadow ? slim : TH_FLT_MAX, st,
                              &tt, &ss, lds_nodes, cached);
    float xt = oswap_f(tt);
    uint32_t xs = oswap_u(ss), xr = oswap_u((uint32_t)r);
    *t = shadow ? xt : tt;
    *slot = shadow ? xs : ss;
    return shadow ? r : (xr != 0u);
  }

  __device__ __forceinline__ float rnd(uint32_t* s) { return th::rand01(s); }

  __device__ v3 random_unit_vector(uint32_t* s) {  // wgsl:202-207
    float z = rnd(s) * 2.0f - 1.0f;
    float a = rnd(s) * 6.2831853f;
    float r = th::fsqrt(th::fmax_(0.0f, 1.0f - z * z));
    float sa, ca;
    th::sincos_(a, &sa, &ca);
    return th::mk(r * ca, r * sa, z);
  }

  // ---------------------------------------------------------------- environment
  __device__ v3 env_texel(int x, int y) {  // wgsl:273-277
    int sw = (int)P.environment_width;
    uint32_t wx = (uint32_t)(((x % sw) + sw) % sw);
    int hm = (int)P.environment_height - 1;
    int cy = y < 0 ? 0 : (y > hm ? hm : y);
    float4 t = S.env[(size_t)cy * P.environment_width + wx];
    return th::mk(t.x, t.y, t.z);
  }
  __device__ v3 env_lookup(const v3& dir) {  // wgsl:280-305
    if (COUNT) lc.env_lookups++;
    v3 ud = th::normalize(dir);
    float u = th::fract(th::fatan2(ud.z, ud.x) / (2.0f * TH_PI) + 0.5f +
                        P.environment_rotation_degrees / 360.0f);
    float v = th::facos(th::clampf(ud.y, -1.0f, 1.0f)) / TH_PI;
    float px = u * (float)P.environment_width - 0.5f;
    float py = v * (float)P.environment_height - 0.5f;
    int bx = (int)th::ffloor(px), by = (int)th::ffloor(py);
    float ax = th::fract(px), ay = th::fract(py);
    v3 top = th::mix(env_texel(bx, by), env_texel(bx + 1, by), ax);
    v3 bottom = th::mix(env_texel(bx, by + 1), env_texel(bx + 1, by + 1), ax);
    return th::muls(th::mix(top, bottom, ay), P.environment_intensity);
  }
  __device__ float env_probability(uint32_t index) {  // wgsl:308-314
    float prev = 0.0f;
    if (S.env_cdf != nullptr) {  // the same values from the compact copy
      if (index > 0u) prev = S.env_cdf[index - 1u];
      return th::fmax_(0.0f, S.env_cdf[index] - prev);
    }
    if (index > 0u) prev = S.env[index - 1u].w;
    return th::fmax_(0.0f, S.env[index].w - prev);
  }
  __device__ float env_pdf(const v3& dir) {  // wgsl:317-335
    v3 ud = th::normalize(dir);
    float u = th::fract(th::fatan2(ud.z, ud.x) / (2.0f * TH_PI) + 0.5f +
                        P.environment_rotation_degrees / 360.0f);
    float theta = th::facos(th::clampf(ud.y, -1.0f, 1.0f));
    float v = theta / TH_PI;
    uint32_t W = P.environment_width, H = P.environment_height;
    uint32_t x = min((uint32_t)th::ffloor(u * (float)W), W - 1u);
    uint32_t y = min((uint32_t)th::ffloor(v * (float)H), H - 1u);
    float prob = env_probability(y * W + x);
    float stt = th::fsin(theta);
    if (prob <= 0.0f || stt <= 0.0f) return 0.0f;
    return prob * (float)(W * H) / (2.0f * TH_PI * TH_PI * stt);
  }
  __device__ EnvSample env_sample(uint32_t* s) {  // wgsl:345-381
    EnvSample out;
    out.dir = th::mk(0, 0, 0);
    out.rad = th::mk(0, 0, 0);
    out.pdf = 0.0f;
    uint32_t W = P.environment_width, H = P.environment_height;
    uint32_t count = W * H;
    if (count == 0u || S.env[count - 1u].w <= 0.0f) return out;
    if (COUNT) lc.env_samples++;
    float sel = th::fmin_(rnd(s), 0.99999994f);
    uint32_t index;
    if (S.env_cdf != nullptr && S.env_rows_w == W) {  // row, then texel within the row
      index = th::env_search_rows(S.env_cdf, S.env_rows, W, H, sel);
    } else {
      uint32_t lo = 0, hi = count;
      while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2u;
        if (S.env[mid].w <= sel)
          lo = mid + 1u;
        else
          hi = mid;
      }
      index = min(lo, count - 1u);
    }
    uint32_t x = index % W, y = index / W;
    float jx = rnd(s);
    float jy = rnd(s);
    jx = th::fmin_(jx, 0.99999994f);
    jy = th::fmin_(jy, 0.99999994f);
    float u = ((float)x + jx) / (float)W;
    float v = ((float)y + jy) / (float)H;
    float theta = TH_PI * v;
    float phi = 2.0f * TH_PI * (u - 0.5f - P.environment_rotation_degrees / 360.0f);
    float stt = th::fsin(theta);
    float sp, cp;
    th::sincos_(phi, &sp, &cp);
    out.dir = th::mk(stt * cp, th::fcos(theta), stt * sp);
    float prob = env_probability(index);
    if (prob <= 0.0f || stt <= 0.0f) return out;
    out.pdf = prob * (float)count / (2.0f * TH_PI * TH_PI * stt);
    out.rad = env_lookup(out.dir);
    return out;
  }

  __device__ static float power_heuristic(float a, float b) {  // wgsl:475-483
    float a2 = a * a, b2 = b * b;
    float den = a2 + b2;
    if (den <= 0.0f) return 0.0f;
    return a2 / den;
  }

  // ---------------------------------------------------------------- lights
  __device__ uint32_t pick_light(uint32_t* s) {  // wgsl:390-403 as a lower-bound search
    // first light whose cumulative weight reaches the target, else the last one: the
    // last light's own comparison cannot change the result, so it is not loaded
    float target = rnd(s) * P.total_light_weight;
    uint32_t n = P.light_count, lo = 0, hi = n - 1u;
    while (lo < hi) {
      uint32_t mid = (lo + hi) >> 1;
      if (S.lights[mid].area_weight_cumulative[2] >= target)
        hi = mid;
      else
        lo = mid + 1u;
    }
    uint32_t idx = lo;
    if (COUNT) lc.lights_scanned += idx + 1u;
    return idx;
  }

  // direct_light (wgsl:406-472) up to its shadow ray: false where the reference returns
  // zero before tracing; otherwise the shadow ray (so, dir, t < limit) and the term the
  // reference returns when that ray is unoccluded.
  __device__ bool light_candidate(const v3& hp, const v3& hn, const v3& albedo, uint32_t* s, v3* so,
                                  v3* dir, float* limit, v3* term) {
    if (P.light_count == 0u || P.total_light_weight <= 0.0f) return false;
    const TH_GpuLight L = load_light(S, pick_light(s));
    v3 pos, ln;
    if (L.kind == 0u) {
      float u = th::fsqrt(rnd(s));
      float v = rnd(s);
      float w0 = 1.0f - u, w1 = u * (1.0f - v), w2 = u * v;
      v3 a = th::mk(L.v0[0], L.v0[1], L.v0[2]);
      v3 b = th::mk(L.v1[0], L.v1[1], L.v1[2]);
      v3 c = th::mk(L.v2[0], L.v2[1], L.v2[2]);
      pos = th::add(th::add(th::smul(w0, a), th::smul(w1, b)), th::smul(w2, c));
      ln = th::normalize(th::cross(th::sub(b, a), th::sub(c, a)));
    } else {
      ln = random_unit_vector(s);
      pos = th::add(th::mk(L.center_radius[0], L.center_radius[1], L.center_radius[2]),
                    th::smul(L.center_radius[3], ln));
    }
    const TH_GpuMaterial em = load_material(S, L.material_index);
    *so = th::add(hp, th::muls(hn, TH_MIN_DISTANCE));
    v3 tl = th::sub(pos, *so);
    float d2 = th::dot(tl, tl);
    if (d2 <= TH_MIN_DISTANCE * TH_MIN_DISTANCE) return false;
    float dist = th::fsqrt(d2);
    *dir = th::divs(tl, dist);
    float sc = th::dot(*dir, hn);
    float lcos = th::dot(th::neg(*dir), ln);
    if (sc <= 0.0f || lcos <= 0.0f) return false;
    *limit = dist - 2.0f * TH_MIN_DISTANCE;
    v3 brdf = th::divs(albedo, TH_PI);
    float area = L.area_weight_cumulative[0];
    float weight = L.area_weight_cumulative[1];
    float sel_pdf = weight / P.total_light_weight;
    float surf_pdf = 1.0f / area;
    float pdf_area = sel_pdf * surf_pdf;
    v3 r = th::mul(brdf, th::mk(em.albedo[0], em.albedo[1], em.albedo[2]));
    r = th::muls(r, em.params[2]);
    r = th::muls(r, sc);
    r = th::muls(r, lcos);
    *term = th::divs(r, d2 * pdf_area);
    return true;
  }

  __device__ v3 direct_light(const Hit& hit, const v3& albedo, uint32_t* s) {  // wgsl:406-472
    v3 so, dir, term;
    float limit;
    if (!light_candidate(hit.p, hit.n, albedo, s, &so, &dir, &limit, &term)) return th::mk(0, 0, 0);
    if (occluded(so, dir, limit)) return th::mk(0, 0, 0);
    return term;
  }

  // direct_environment (wgsl:486-505) up to its shadow ray: the environment sample and
  // its cosine; false where the reference returns zero before tracing.
  __device__ bool env_candidate(const v3& hn, uint32_t* s, EnvSample* es, float* sc) {
    *es = env_sample(s);
    *sc = th::fmax_(0.0f, th::dot(hn, es->dir));
    return !(*sc <= 0.0f || es->pdf <= 0.0f);
  }
  // The unoccluded term of direct_environment for a given MIS weight.
  __device__ static v3 env_term(const v3& albedo, const EnvSample& es, float sc, float mis) {
    v3 brdf = th::divs(albedo, TH_PI);
    v3 r = th::mul(brdf, es.rad);
    r = th::muls(r, sc);
    r = th::muls(r, mis);
    return th::divs(r, es.pdf);
  }

  __device__ v3 direct_environment(const Hit& hit, const v3& albedo, uint32_t* s) {  // wgsl:486-505
    if (P.background_kind != 2u) return th::mk(0, 0, 0);
    EnvSample es;
    float sc;
    if (!env_candidate(hit.n, s, &es, &sc)) return th::mk(0, 0, 0);
    // blocked when the closest hit is within MAX_DISTANCE (t <= 10000 <=> t < next float)
    if (occluded(th::add(hit.p, th::muls(hit.n, TH_MIN_DISTANCE)), es.dir, 10000.001f))
      return th::mk(0, 0, 0);
    return env_term(albedo, es, sc, power_heuristic(es.pdf, sc / TH_PI));
  }

  // ---------------------------------------------------------------- textures
  __device__ v3 tex_texel(const TH_GpuMaterial& m, int x, int y) {  // wgsl:233-244
    int sw = (int)m.texture_width, sh = (int)m.texture_height;
    uint32_t wx = (uint32_t)(((x % sw) + sw) % sw), wy = (uint32_t)(((y % sh) + sh) % sh);
    uint32_t packed = S.tex[m.texture_offset + wy * m.texture_width + wx];
    return th::mk(srgb_lut[packed & 255u], srgb_lut[(packed >> 8u) & 255u],
                  srgb_lut[(packed >> 16u) & 255u]);
  }
  __device__ v3 base_color(const TH_GpuMaterial& m, const Hit& h) {  // wgsl:247-270
    v3 albedo = th::mk(m.albedo[0], m.albedo[1], m.albedo[2]);
    if (m.texture_width == 0u || m.texture_height == 0u) return albedo;
    float ru = th::fract(h.tu), rv = th::fract(h.tv);
    float px = ru * (float)m.texture_width - 0.5f, py = rv * (float)m.texture_height - 0.5f;
    int bx = (int)th::ffloor(px), by = (int)th::ffloor(py);
    float ax = th::fract(px), ay = th::fract(py);
    v3 top = th::mix(tex_texel(m, bx, by), tex_texel(m, bx + 1, by), ax);
    v3 bottom = th::mix(tex_texel(m, bx, by + 1), tex_texel(m, bx + 1, by + 1), ax);
    return th::mul(albedo, th::mix(top, bottom, ay));
  }

  // ---------------------------------------------------------------- path
  __device__ static v3 reflect_vec(const v3& d, const v3& n) {  // wgsl:919-921
    return th::sub(d, th::muls(n, th::dot(n, d) * 2.0f));
  }
  __device__ static v3 refract_vec(const v3& d, const v3& n, float ratio) {  // wgsl:924-929
    float ct = th::fmin_(th::dot(th::neg(d), n), 1.0f);
    v3 perp = th::smul(ratio, th::add(d, th::smul(ct, n)));
    v3 par = th::smul(-th::fsqrt(th::fabs_(1.0f - th::dot(perp, perp))), n);
    return th::add(perp, par);
  }
  __device__ static float reflectance(float cosine, float ratio) {  // wgsl:932-937
    float r0 = (1.0f - ratio) / (1.0f + ratio);
    r0 = r0 * r0;
    float t = 1.0f - cosine;
    return r0 + (1.0f - r0) * t * t * t * t * t;
  }

  __device__ v3 ray_color(v3 o, v3 d, uint32_t* s) {  // wgsl:805-916
    v3 throughput = th::mk(1, 1, 1);
    v3 radiance = th::mk(0, 0, 0);
    bool include_emissive = true, prev_diffuse = false;
    float prev_pdf = 0.0f;
    for (uint32_t i = 0; i < P.max_bounces; ++i) {
      Hit rec;
      if (!trace(o, d, &rec)) {
        if (P.background_kind == 0u) {
          float lt = (th::normalize(d).y + 1.0f) * 0.5f;
          radiance = th::add(radiance, th::mul(throughput, th::mix(th::mk(1.0f, 1.0f, 1.0f),
                                                                    th::mk(0.35f, 0.65f, 1.0f), lt)));
        } else if (P.background_kind == 2u) {
          float mis = 1.0f;
          if (prev_diffuse) mis = power_heuristic(prev_pdf, env_pdf(d));
          radiance = th::add(radiance, th::muls(th::mul(throughput, env_lookup(d)), mis));
        }
        break;
      }
      const TH_GpuMaterial m = load_material(S, rec.mat);
      v3 att;
      v3 nd;
      bool brk = false;
      if (m.kind == 0u) {
        v3 albedo = base_color(m, rec);
        radiance = th::add(radiance, th::mul(throughput, direct_light(rec, albedo, s)));
        radiance = th::add(radiance, th::mul(throughput, direct_environment(rec, albedo, s)));
        v3 dir = th::add(rec.n, random_unit_vector(s));
        if (th::dot(dir, dir) < TH_EPSILON) dir = rec.n;
        nd = th::normalize(dir);
        att = albedo;
        include_emissive = false;
        prev_diffuse = true;
        prev_pdf = th::fmax_(0.0f, th::dot(rec.n, nd)) / TH_PI;
      } else if (m.kind == 1u) {
        v3 refl = reflect_vec(d, rec.n);
        float rough = th::clampf(m.params[0], 0.0f, 1.0f);
        v3 dir = th::add(refl, th::smul(rough, random_unit_vector(s)));
        if (th::dot(dir, rec.n) <= 0.0f) brk = true;
        nd = th::normalize(dir);
        att = th::mk(m.albedo[0], m.albedo[1], m.albedo[2]);
        include_emissive = true;
        prev_diffuse = false;
      } else if (m.kind == 2u) {
        float ratio = rec.front ? 1.0f / m.params[1] : m.params[1];
        v3 ud = th::normalize(d);
        float ct = th::fmin_(1.0f, th::dot(th::neg(ud), rec.n));
        float stt = th::fsqrt(1.0f - ct * ct);
        bool cannot = (ratio * stt) > 1.0f;
        v3 dir;
        if (cannot || reflectance(ct, ratio) > rnd(s))
          dir = reflect_vec(ud, rec.n);
        else
          dir = refract_vec(ud, rec.n, ratio);
        nd = th::normalize(dir);
        att = th::mk(1, 1, 1);
        include_emissive = true;
        prev_diffuse = false;
      } else {
        if (m.kind == 3u && include_emissive)
          radiance = th::add(radiance, th::muls(th::mul(throughput, th::mk(m.albedo[0], m.albedo[1],
                                                                            m.albedo[2])),
                                                m.params[2]));
        brk = true;
      }
      if (brk) break;
      throughput = th::mul(throughput, att);
      o = rec.p;
      d = nd;
    }
    return radiance;
  }
};

// ray_color (wgsl:805-916) for an octet of lanes.  The shadow rays of a diffuse vertex
// are traced by the partner quad while the path's quad traces the next bounce, and their
// terms are added when both are back, before anything the next vertex adds.  The
// continuation ray does not depend on the shadow rays' outcome, so the draws, terms and
// additions are those of ray_color in the same order and the result is identical.  The
// shadow rays of a sample's last vertex go with the next sample's primary ray; a sample's
// radiance is added to the pixel sum once it is complete, in sample order.
// ENV false: scenes whose background is not an environment map (no environment shadow
// rays).
template <bool ENV, int TMODE, bool COUNT>
struct OctetPath {
  Integrator<TMODE, COUNT>& it;
  bool shadow;  // this lane's quad traces shadow rays
  // the last diffuse vertex: its throughput, shadow origin, light and environment
  // shadow rays (pl / pe: ray pending) and their unoccluded terms
  bool pd = false, pl = false, pe = false;
  v3 pthr, so, ldir, edir, lterm, eterm;
  float llim = 0.0f;
  // a finished sample's radiance waiting for its last vertex's shadow rays
  bool carry = false;
  v3 crad;
  float cr = 0.0f, cg = 0.0f, cb = 0.0f, ca = 0.0f;  // pixel sums

  __device__ void emit(const v3& L) {
    cr = cr + L.x;
    cg = cg + L.y;
    cb = cb + L.z;
    ca = ca + 1.0f;
  }

  // the current sample: ray, throughput, radiance so far, bounce index
  v3 o, d, throughput, radiance;
  bool include_emissive = true, prev_diffuse = false;
  float prev_pdf = 0.0f;
  uint32_t i = 0;

  // Starts a sample from the camera ray (o, d).
  __device__ __forceinline__ void begin(const v3& o0, const v3& d0) {
    o = o0;
    d = d0;
    throughput = th::mk(1, 1, 1);
    radiance = th::mk(0, 0, 0);
    include_emissive = true;
    prev_diffuse = false;
    prev_pdf = 0.0f;
    i = 0;
  }

  // Drops the current sample and any pending shadow rays (the pixel sums are kept).
  __device__ void abandon() { pd = pl = pe = carry = false; }

  // One sample from the camera ray (o, d); `last`: no sample follows in this launch.
  __device__ __forceinline__ void sample(const v3& o0, const v3& d0, uint32_t* s, bool last) {
    begin(o0, d0);
    while (step(s, last)) {
    }
  }

  // One vertex of the current sample (the extension ray and the previous vertex's
  // shadow rays in one traversal, then shading); false once the sample has ended (its
  // radiance emitted, or carried to the next sample when `last` is false).
  __device__ __forceinline__ bool step(uint32_t* s, bool last) {
    const DevScene& S = it.S;
    const TH_GpuRenderParams& P = it.P;
    const v3 zero = th::mk(0, 0, 0);
    bool more = false, emits = true;
    do {  // one vertex; `break` ends the sample
      bool ext = i < P.max_bounces;
      if (!ext && !pd) break;
      if (!ext && !last) {  // trace this vertex's shadow rays with the next primary ray
        carry = true;
        crad = radiance;
        emits = false;
        break;
      }
      // the shadow quad takes the light's shadow ray, else the environment's; a second
      // step traces the environment's when both are pending
      float t;
      uint32_t slot;
      bool first_env = ENV && !pl;
      bool occ1 = it.octet_step(shadow, ext, o, d, pl || (ENV && pe), so, first_env ? edir : ldir,
                                first_env ? 10000.001f : llim, &t, &slot);
      bool occ_l = !first_env && occ1, occ_e = first_env && occ1;
      if (ENV && pl && pe) {
        float t2;
        uint32_t s2;
        occ_e = it.octet_step(shadow, false, o, d, true, so, edir, 10000.001f, &t2, &s2);
      }
      if (pd) {
        v3 r = carry ? crad : radiance;
        r = th::add(r, th::mul(pthr, occ_l ? zero : lterm));
        r = th::add(r, th::mul(pthr, (ENV && !occ_e) ? eterm : zero));
        if (carry) emit(r);
        else radiance = r;
        pd = pl = pe = carry = false;
      }
      if (!ext) break;
      if (slot == THD_NO_SLOT || t > TH_MAX_DISTANCE) {
        if (P.background_kind == 0u) {
          float lt = (th::normalize(d).y + 1.0f) * 0.5f;
          radiance = th::add(radiance, th::mul(throughput, th::mix(th::mk(1.0f, 1.0f, 1.0f),
                                                                    th::mk(0.35f, 0.65f, 1.0f), lt)));
        } else if (ENV) {
          float mis = 1.0f;
          if (prev_diffuse) mis = it.power_heuristic(prev_pdf, it.env_pdf(d));
          radiance = th::add(radiance, th::muls(th::mul(throughput, it.env_lookup(d)), mis));
        }
        break;
      }
      Hit rec;
      finalize_slot(S, slot, o, d, t, &rec);
      const TH_GpuMaterial m = load_material(S, rec.mat);
      v3 att, nd;
      bool brk = false;
      if (m.kind == 0u) {
        v3 albedo = it.base_color(m, rec);
        pd = true;
        pthr = throughput;
        so = th::add(rec.p, th::muls(rec.n, TH_MIN_DISTANCE));
        lterm = zero;
        pl = it.light_candidate(rec.p, rec.n, albedo, s, &so, &ldir, &llim, &lterm);
        eterm = zero;
        if (ENV) {
          EnvSample es;
          float sc;
          pe = it.env_candidate(rec.n, s, &es, &sc);
          if (pe) {
            edir = es.dir;
            eterm = it.env_term(albedo, es, sc, it.power_heuristic(es.pdf, sc / TH_PI));
          }
        }
        v3 dir = th::add(rec.n, it.random_unit_vector(s));
        if (th::dot(dir, dir) < TH_EPSILON) dir = rec.n;
        nd = th::normalize(dir);
        att = albedo;
        include_emissive = false;
        prev_diffuse = true;
        prev_pdf = th::fmax_(0.0f, th::dot(rec.n, nd)) / TH_PI;
      } else if (m.kind == 1u) {
        v3 refl = it.reflect_vec(d, rec.n);
        float rough = th::clampf(m.params[0], 0.0f, 1.0f);
        v3 dir = th::add(refl, th::smul(rough, it.random_unit_vector(s)));
        if (th::dot(dir, rec.n) <= 0.0f) brk = true;
        nd = th::normalize(dir);
        att = th::mk(m.albedo[0], m.albedo[1], m.albedo[2]);
        include_emissive = true;
        prev_diffuse = false;
      } else if (m.kind == 2u) {
        float ratio = rec.front ? 1.0f / m.params[1] : m.params[1];
        v3 ud = th::normalize(d);
        float ct = th::fmin_(1.0f, th::dot(th::neg(ud), rec.n));
        float stt = th::fsqrt(1.0f - ct * ct);
        bool cannot = (ratio * stt) > 1.0f;
        v3 dir;
        if (cannot || it.reflectance(ct, ratio) > it.rnd(s))
          dir = it.reflect_vec(ud, rec.n);
        else
          dir = it.refract_vec(ud, rec.n, ratio);
        nd = th::normalize(dir);
        att = th::mk(1, 1, 1);
        include_emissive = true;
        prev_diffuse = false;
      } else {
        if (m.kind == 3u && include_emissive)
          radiance = th::add(radiance, th::muls(th::mul(throughput, th::mk(m.albedo[0], m.albedo[1],
                                                                            m.albedo[2])),
                                                m.params[2]));
        brk = true;
      }
      if (brk) break;
      throughput = th::mul(throughput, att);
      o = rec.p;
      d = nd;
      ++i;
      more = true;
    } while (false);
    if (!more && emits) emit(radiance);
    return more;
  }
};

// Primary ray of wgsl:525-534 (two draws: x jitter, then y jitter).
__device__ __forceinline__ void camera_ray(const TH_GpuCamera& cam, const TH_GpuRenderParams& P,
                                           uint32_t x, uint32_t y, uint32_t* s, v3* o, v3* d) {
  float u = ((float)x + th::rand01(s)) / (float)P.width;
  float v = 1.0f - (((float)y + th::rand01(s)) / (float)P.height);
  *o = th::mk(cam.origin[0], cam.origin[1], cam.origin[2]);
  v3 dd = th::add(
      th::add(th::mk(cam.lower_left_corner[0], cam.lower_left_corner[1], cam.lower_left_corner[2]),
              th::smul(u, th::mk(cam.horizontal[0], cam.horizontal[1], cam.horizontal[2]))),
      th::smul(v, th::mk(cam.vertical[0], cam.vertical[1], cam.vertical[2])));
  *d = th::normalize(th::sub(dd, *o));
}

}  // namespace thd
