// kg_math.hpp -- f32 helpers shared by the HIP kernels and the host runtime.
//
// Every function here performs the same IEEE-754 single-precision operations,
// in the same order, as the krABMaga code it restates, so host and device
// produce identical bits (the library is compiled with -ffp-contract=off and
// correctly rounded f32 division and square root).
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define KG_HD __host__ __device__ __forceinline__
#else
// Host-only builds (the CPU driver of the kernel phases used by the tests).
#define KG_HD inline
#endif

namespace kg {

// Philox4x32-10 (Salmon et al., SC'11): multipliers and Weyl increments as
// published there.
constexpr uint32_t kPhiloxM0 = 0xD2511F53u;
constexpr uint32_t kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u;
constexpr uint32_t kPhiloxW1 = 0xBB67AE85u;
constexpr uint32_t kTagStep = 0x53544550u;  // "STEP"
constexpr uint32_t kTagInit = 0x494E4954u;  // "INIT"

struct u32x4 {
  uint32_t v[4];
};

KG_HD u32x4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                          uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += kPhiloxW0;
      k1 += kPhiloxW1;
    }
    // one 32 x 32 -> 64 multiply gives both halves (a single v_mad_u64_u32)
    const uint64_t p0 = (uint64_t)kPhiloxM0 * c0, p1 = (uint64_t)kPhiloxM1 * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c1 ^ k0;
    const uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  u32x4 o;
  o.v[0] = c0;
  o.v[1] = c1;
  o.v[2] = c2;
  o.v[3] = c3;
  return o;
}

// u32 -> [0,1): the top 24 bits times 2^-24 (rand's Standard f32 mapping).
KG_HD float u01(uint32_t u) { return (float)(u >> 8) * (1.0f / 16777216.0f); }

// The two uniforms Bird::step draws (tests/model/flockers/bird.rs:113-117),
// from the counter (id, step, STEP) under the run's 64-bit seed.
KG_HD void step_uniforms(uint64_t seed, uint32_t id, uint64_t step, float* r1, float* r2) {
  u32x4 o = philox4x32_10(id, (uint32_t)step, (uint32_t)(step >> 32), kTagStep, (uint32_t)seed,
                          (uint32_t)(seed >> 32));
  *r1 = u01(o.v[0]);
  *r2 = u01(o.v[1]);
}

// Initial state of agent `id` (tests/model/flockers/state.rs:42-47, Philox
// in place of rand::rng()): pos = dim * U[0,1); heading (0,0) or a unit
// vector at a uniform angle, drawn by rejection from [-1,1)^2.
KG_HD void init_agent(uint32_t id, float w, float h, uint64_t seed, int heading_mode, float* x,
                      float* y, float* hx, float* hy) {
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  u32x4 o = philox4x32_10(id, 0u, 0u, kTagInit, k0, k1);
  *x = w * u01(o.v[0]);
  *y = h * u01(o.v[1]);
  if (heading_mode == 0) {
    *hx = 0.0f;
    *hy = 0.0f;
    return;
  }
  float u = 2.0f * u01(o.v[2]) - 1.0f;
  float v = 2.0f * u01(o.v[3]) - 1.0f;
  float s = u * u + v * v;
  uint32_t k = 0;
  while (!(s > 0.0f && s <= 1.0f)) {
    o = philox4x32_10(id, ++k, 0u, kTagInit, k0, k1);
    u = 2.0f * u01(o.v[2]) - 1.0f;
    v = 2.0f * u01(o.v[3]) - 1.0f;
    s = u * u + v * v;
  }
  const float r = sqrtf(s);
  *hx = u / r;
  *hy = v / r;
}

KG_HD int32_t imin(int32_t a, int32_t b) { return a < b ? a : b; }
KG_HD int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }

// Rust `f as i32`: saturating, NaN -> 0.  Branch-free: the clamp keeps the
// conversion in range (fmaxf of NaN gives the bound, fixed by the last
// select); 2147483520 is the largest float below 2^31.
KG_HD int32_t f2i(float f) {
  const float c = fminf(fmaxf(f, -2147483648.0f), 2147483520.0f);
  int32_t r = (int32_t)c;
  r = f >= 2147483648.0f ? INT32_MAX : r;
  return f != f ? 0 : r;
}

// src/engine/fields/field_2d.rs:926-932
KG_HD int32_t t_transform(int32_t n, int32_t size) {
  if (n >= 0) return n % size;
  return (n % size) + size;
}

// General case of toroidal_transform, kept out of line: fmodf expands to a
// long reduction loop and almost never runs.
#ifdef __HIPCC__
__host__ __device__ inline __attribute__((noinline))
#else
inline
#endif
float toroidal_transform_slow(float val, float dim) {
  float v = fmodf(val, dim);
  if (v < 0.0f) v += dim;
  return v;
}

// src/engine/fields/field_2d.rs:1004-1014 (Rust `%` on f32 is fmod).  For
// dim <= val < 2 dim, fmod(val, dim) = val - dim and the subtraction is exact
// (Sterbenz); for -dim < val < 0, fmod(val, dim) = val.  Both shortcuts give
// the bits of the general formula; other values take the fmodf path.
KG_HD float toroidal_transform(float val, float dim) {
  if (val >= 0.0f && val < dim) return val;
  if (dim > 0.0f) {
    if (val >= dim && val < 2.0f * dim) return val - dim;
    if (val < 0.0f && val > -dim) return val + dim;
  }
  return toroidal_transform_slow(val, dim);
}

// src/engine/fields/field_2d.rs:988-1002.  dim * 0.5f is dim / 2.0 to the
// bit (the same real value, correctly rounded), without a division.
KG_HD float toroidal_distance(float val1, float val2, float dim) {
  if (fabsf(val1 - val2) <= dim * 0.5f) return val1 - val2;
  const float d = toroidal_transform(val1, dim) - toroidal_transform(val2, dim);
  if (d * 2.0f > dim) return d - dim;
  if (d * 2.0f < -dim) return d + dim;
  return d;
}

// Squared distance exactly as field_2d.rs:974-986 forms it before its
// sqrt: fl(fl(dx*dx) + fl(dy*dy)).  Because f32 sqrt is correctly rounded
// and monotone, sqrt(s) <= r holds exactly when s <= sq_threshold(r), so the
// kernels compare squared distances and never take a square root.
KG_HD float dist2(float x1, float y1, float x2, float y2, float w, float h, bool tor) {
  float dx, dy;
  if (tor) {
    dx = toroidal_distance(x1, x2, w);
    dy = toroidal_distance(y1, y2, h);
  } else {
    dx = x1 - x2;
    dy = y1 - y2;
  }
  return dx * dx + dy * dy;
}

// Field geometry, derived on the host from Field2D::new's arguments
// (field_2d.rs:304-322) plus the device binning parameters.
//
// Device bins: the bag grid (dw x dh cells) is cut into tiles of tx columns
// by ty = tr * hb rows; a tile holds tx * tr bins, each bin a strip of hb
// cells of one column.  Bins are numbered tile-major (tiles row-major, then
// column inside the tile, then row block), so one tile's agents form one
// contiguous range of the read buffer.  tx, tr and hb are powers of two.
struct Geom {
  float w, h, d;         // width, height, discretization
  int32_t tor;           // toroidal flag
  int32_t max_x, max_y;  // ceil(w/d), ceil(h/d)  (field_2d.rs:403-404)
  int32_t dw, dh;        // max_x + 1, max_y + 1  (bag grid incl. padding)
  int32_t tx_log2, tr_log2, hb_log2;
  int32_t tx, tr, hb, ty;  // tile columns, row blocks per tile, cells per block, tile rows
  int32_t ntx, nty;        // tiles per row / column of the tile grid
  uint32_t ntiles;
  uint32_t nbins;          // ntiles * tx * tr
  float inv_d;             // fl(1 / d), for disc_axis
};

// floor(v / d) as Rust's `(v / d).floor() as i32` computes it: the correctly
// rounded quotient, floored, converted with saturation.  Fast path: with
// r = fl(1/d), q = fl(v * r) is within 3 * 2^-24 |v/d| of the quotient, and
// fl(v/d) within 2^-24 |v/d|, so they differ by less than eps = 2^-21 |q|;
// when q lies at least eps inside an integer interval [k, k + 1) the
// quotient lies in it too and its floor is k.  Otherwise (near a cell
// boundary, |q| >= 2^22, tiny, zero or not finite) the division itself is
// taken.  Checked against floor(v / d) over every float in [0, 2^15) and
// [-64, 0), the boundaries of 2*10^7 cells and special values, for eleven
// discretizations including 10/1.5.
KG_HD int32_t disc_axis(float v, float d, float inv_d) {
  const float q = v * inv_d;
  const float k = floorf(q);
  const float a = fabsf(q);
  const float eps = a * 0x1p-21f;
  if (a >= 0x1p-100f && a < 0x1p22f && q - k >= eps && (k + 1.0f) - q >= eps)
    return (int32_t)k;
  return f2i(floorf(v / d));
}

// field_2d.rs:328-339
KG_HD void discretize(const Geom& g, float x, float y, int32_t* cx, int32_t* cy) {
  *cx = disc_axis(x, g.d, g.inv_d);
  *cy = disc_axis(y, g.d, g.inv_d);
}

}  // namespace kg
