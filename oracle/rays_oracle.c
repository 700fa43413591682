/*
 * CPU oracle of the path tracer's ray generation -- TEST INFRASTRUCTURE ONLY.
 *
 * Restates, op for op, the rays the HIP library generates, which follow the
 * reference's shaders (marigold/assets/shaders/slang):
 *   PCG stream ............... lib/random.slang:5-27
 *   unit / cosine vectors .... lib/random.slang:29-41
 *   frame jitter ............. pathtrace.slang:46-61
 *   primary direction ........ lib/screen.slang:48-72, pathtrace.slang:61-66
 *   bounce origin ............ lib/intersect.slang:49-51 (hit + 0.001 n)
 * Only tests/ load this file's library, and only as the checker.
 *
 * Rays are 8 floats: org[3], 0, dir[3], tmax (+inf), as mgbvh_ray.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

/* OR_PLANT (shared with bvh_oracle.c) selects a deliberately wrong variant
 * in the libraries the tests build for that purpose:
 *   6: clip y not flipped (ty * 2 - 1 instead of 1 - ty * 2, screen.slang:51)
 *   7: pixel centres (0.5 added) instead of the frame jitter */
#ifndef OR_PLANT
#define OR_PLANT 0
#endif

/* random.slang:5-9 */
uint32_t or_pcg(uint32_t* seed) {
  uint32_t state = *seed * 747796405u + 2891336453u;
  uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  *seed = (word >> 22u) ^ word;
  return *seed;
}

/* random.slang:24-27: float(getUint()) / float(0xffffffff), both conversions
 * rounded to nearest */
float or_pcg_float(uint32_t* seed) {
  uint32_t u = or_pcg(seed);
  return (float)u / (float)0xffffffffu;
}

/* pathtrace.slang:50-61: Random(frameCount); jitter = 2 draws - 0.5, applied
 * when the image accumulates */
void or_frame_jitter(uint32_t frame_count, int accumulate, float* j2) {
  uint32_t s = frame_count;
  float j0 = or_pcg_float(&s) - 0.5f;
  float j1 = or_pcg_float(&s) - 0.5f;
  j2[0] = accumulate ? j0 : 0.0f;
  j2[1] = accumulate ? j1 : 0.0f;
}

static void put_ray(float* r, const float* o, const float* d) {
  r[0] = o[0]; r[1] = o[1]; r[2] = o[2];
  r[3] = 0.0f;
  r[4] = d[0]; r[5] = d[1]; r[6] = d[2];
  r[7] = INFINITY;
}

/* Primary rays, row-major from the top-left pixel.  m16: the inverse
 * view-projection matrix, column-major.  texcoord = (pixel + jitter) / size,
 * clip = (2 tx - 1, 1 - 2 ty, 1), t = M (clip, 1) (sums left to right),
 * world = t.xyz / t.w, dir = normalize(world - pos) = v / sqrt((x x + y y) + z z). */
void or_gen_view_rays(const float* m16, const float* pos, uint32_t w, uint32_t h, float jx,
                      float jy, float* rays8) {
  for (uint32_t y = 0; y < h; ++y)
    for (uint32_t x = 0; x < w; ++x) {
      float px = (float)x + jx, py = (float)y + jy;
      if (OR_PLANT == 7) { px = (float)x + 0.5f; py = (float)y + 0.5f; }
      float tx = px / (float)w, ty = py / (float)h;
      float cl[3];
      cl[0] = tx * 2.0f - 1.0f;
      cl[1] = OR_PLANT == 6 ? ty * 2.0f - 1.0f : 1.0f - ty * 2.0f;
      cl[2] = 1.0f;
      float t[4];
      for (int i = 0; i < 4; ++i) {
        float acc = m16[i] * cl[0] + m16[4 + i] * cl[1];
        acc = acc + m16[8 + i] * cl[2];
        t[i] = acc + m16[12 + i];
      }
      float d[3];
      for (int a = 0; a < 3; ++a) d[a] = t[a] / t[3] - pos[a];
      float l2 = d[0] * d[0] + d[1] * d[1];
      float l = sqrtf(l2 + d[2] * d[2]);
      for (int a = 0; a < 3; ++a) d[a] = d[a] / l;
      put_ray(rays8 + 8 * ((size_t)y * w + x), pos, d);
    }
}

static void normalize3(float* v) {
  float l2 = v[0] * v[0] + v[1] * v[1];
  float l = sqrtf(l2 + v[2] * v[2]);
  v[0] /= l; v[1] /= l; v[2] /= l;
}

/* One bounce per ray, seeded with (ray index + seed) as pathtrace.slang:49
 * seeds pixelId + frameOffset.  getUnitVector (random.slang:29-36): two draws
 * u0, u1; phi = u0 * 2 pi; z = 2 u1 - 1; (sin phi, cos phi) * sqrt(1 - z z), z.
 * A hit: the triangle's geometric normal cross(v2 - v1, v3 - v1) (taken to
 * world space by the transpose of world_to_object's linear part for a scene
 * pass), normalised and turned against the ray; origin hit point + 0.001 n;
 * direction normalize(n + unit vector) (getCosineVector, random.slang:38-41).
 * A miss: origin kept, direction the unit vector.
 * hits4: {t bits, prim, inst, tri_slot}; inst16: per instance world_to_object
 * (12 floats) then the BLAS index; blas_voff: vertex offset per BLAS.  inst16
 * NULL for a single mesh.  The sine and cosine are libm's, so directions agree
 * with the GPU's to a few ulps, not bit for bit. */
void or_gen_bounce(const float* rays8, const uint32_t* hits4, uint32_t n, const float* inst16,
                   const uint32_t* blas_voff, const float* pos, uint32_t vs, const uint32_t* tris,
                   uint32_t seed, float* out8) {
  for (uint32_t i = 0; i < n; ++i) {
    const float* q = rays8 + 8 * (size_t)i;
    const uint32_t* hh = hits4 + 4 * (size_t)i;
    uint32_t s = i + seed;
    float xy0 = or_pcg_float(&s) * 6.283185307179586f;
    float xy1 = 2.0f * or_pcg_float(&s) - 1.0f;
    float rr = sqrtf(fmaxf(0.0f, 1.0f - xy1 * xy1));
    float uvec[3] = {sinf(xy0) * rr, cosf(xy0) * rr, xy1};
    float o[3], d[3];
    if (hh[1] != 0xffffffffu) {
      float t;
      memcpy(&t, &hh[0], 4);
      uint32_t voff = 0;
      const float* m = NULL;
      if (inst16) {
        m = inst16 + 16 * (size_t)hh[2];
        uint32_t b;
        memcpy(&b, m + 12, 4);
        voff = blas_voff[b];
      }
      const uint32_t* tr = tris + 3 * (size_t)hh[3];
      const float* v1 = pos + (size_t)(tr[0] + voff) * vs;
      const float* v2 = pos + (size_t)(tr[1] + voff) * vs;
      const float* v3 = pos + (size_t)(tr[2] + voff) * vs;
      float e1[3], e2[3], nl[3], nrm[3];
      for (int a = 0; a < 3; ++a) { e1[a] = v2[a] - v1[a]; e2[a] = v3[a] - v1[a]; }
      nl[0] = e1[1] * e2[2] - e1[2] * e2[1];
      nl[1] = e1[2] * e2[0] - e1[0] * e2[2];
      nl[2] = e1[0] * e2[1] - e1[1] * e2[0];
      if (m) {
        for (int a = 0; a < 3; ++a) {
          float acc = m[a] * nl[0] + m[4 + a] * nl[1];
          nrm[a] = acc + m[8 + a] * nl[2];
        }
      } else {
        for (int a = 0; a < 3; ++a) nrm[a] = nl[a];
      }
      normalize3(nrm);
      float dn = nrm[0] * q[4] + nrm[1] * q[5];
      if (dn + nrm[2] * q[6] > 0.0f)
        for (int a = 0; a < 3; ++a) nrm[a] = -nrm[a];
      for (int a = 0; a < 3; ++a) {
        float p = q[a] + t * q[4 + a];
        o[a] = p + 0.001f * nrm[a];
        d[a] = nrm[a] + uvec[a];
      }
      normalize3(d);
    } else {
      for (int a = 0; a < 3; ++a) { o[a] = q[a]; d[a] = uvec[a]; }
    }
    put_ray(out8 + 8 * (size_t)i, o, d);
  }
}
