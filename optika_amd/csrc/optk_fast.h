// optk_fast.h -- the float32 form of the sequential trace (OPTK_F32).
//
// Same physics as optk_surface.h (AbstractSurface.propagate_rays,
// optika/surfaces.py:123-198), arranged for throughput instead of
// bit-for-bit agreement with numpy:
//
//  * Rays enter and leave in float32 (half the HBM bytes of float64).
//  * Each surface works in its own sag frame.  The move between two surfaces
//    is ONE precomposed affine map (previous surface frame -> this sag frame)
//    and the ray's position and direction are carried in float64 across it:
//    this is the only place where coordinates are large (hundreds of mm), and
//    CDNA4 issues float64 add/multiply/FMA at the rate of unpacked float32,
//    so float64 there costs registers, not throughput.
//  * The intercept is solved in float32 in sag-frame coordinates relative to
//    the ray's crossing of the vertex plane (z = 0), where every quantity is
//    small.  The residual sag is written without cancellation (the toroid's
//    r - sqrt(q) becomes (r^2 - q) / (r + sqrt(q)); conic roots use the
//    cancellation-free quadratic formula).  Surfaces flagged steep take two
//    more Newton steps of the residual in float64.
//  * Apertures, material, ruling and multilayer efficiencies are float32;
//    normals are float64 (fast square roots) at the float64 hit point.  The multilayer transfer matrices run both polarizations in
//    one pass and evaluate each layer's index, cosine and phase once (Snell
//    invariant n sin(theta)); below a grazing threshold, or where the
//    reference's |U11| < 1e10 guard trips, the float64 routine of
//    optk_physics.h is used instead.
//  * Snell's law and the grating equation update the float64 direction with
//    fast float64 reciprocals and square roots (hardware seed + Newton).
//
// Accuracy is measured, not assumed: the GPU tests state the RMS spot and
// direction error of this form against the float64 form.
#pragma once

#include <string.h>

#include "optk_surface.h"

namespace optk {

// ---------------------------------------------------------------------------
// fast float64 reciprocal, square root and reciprocal square root: the
// hardware approximations (relative error below 2^-24 on gfx950, measured
// over 2^20 arguments spanning 64 binades) refined by one Newton step to
// about 2^-48 relative -- far below what the float32 form's outputs resolve
// (its float64 directions need ~1e-11).  Not correctly rounded.  Zeros,
// infinities and NaNs pass through as IEEE division and sqrt would give them.
// ---------------------------------------------------------------------------

OPTK_HD double rcp64(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  double r1 = fma(r, e, r);
  return (e == e) ? r1 : r;  // x = 0 or inf: the seed is exact
#else
  return 1.0 / x;
#endif
}

// the same for x known finite and nonzero (no special-value select)
OPTK_HD double rcp64_finite(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(x);
  return fma(r, fma(-x, r, 1.0), r);
#else
  return 1.0 / x;
#endif
}

// reciprocal square root of x known finite and positive
OPTK_HD double rsqrt64_finite(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  double y = __builtin_amdgcn_rsq(x);
  return fma(0.5 * y, fma(-x * y, y, 1.0), y);
#else
  return 1.0 / sqrt(x);
#endif
}

// square root from the refined reciprocal square root and one more Newton
// step on x - g^2; zeros and +inf are returned as they are (one class test),
// negative and NaN arguments give NaN through the hardware seed
OPTK_HD double sqrt64(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  double y = rsqrt64_finite(x);
  double g = x * y;
  double d = fma(-g, g, x);
  g = fma(d, 0.5 * y, g);
  // class mask: -0 (bit 5), +0 (bit 6), +inf (bit 9)
  return __builtin_amdgcn_class(x, (1 << 5) | (1 << 6) | (1 << 9)) ? x : g;
#else
  return sqrt(x);
#endif
}

// ---------------------------------------------------------------------------
// float32 complex helpers (magnitudes in the transfer matrices are O(1):
// plain |b|^2 division, no Smith scaling)
// ---------------------------------------------------------------------------

using Cf = C<float>;

OPTK_HD Cf fcdiv(Cf a, Cf b) {
  float inv = 1.0f / (b.re * b.re + b.im * b.im);
  return {(a.re * b.re + a.im * b.im) * inv, (a.im * b.re - a.re * b.im) * inv};
}

OPTK_HD Cf fcrcp(Cf b) {
  float inv = 1.0f / (b.re * b.re + b.im * b.im);
  return {b.re * inv, -b.im * inv};
}

OPTK_HD Cf fcsqrt(Cf z) {  // principal branch
  float x = z.re, y = z.im;
  float d = sqrtf(x * x + y * y);
  if (x >= 0.0f) {
    float r = sqrtf(0.5f * (d + x));
    float s = (r > 0.0f) ? 0.5f * y / r : 0.0f;
    return {r, s};
  }
  float s = sqrtf(0.5f * (d - x));
  float r = 0.5f * fabsf(y) / s;
  return {r, copysignf(s, y)};
}

OPTK_HD float fabs2(Cf a) { return a.re * a.re + a.im * a.im; }

// ---------------------------------------------------------------------------
// records of the float32 form, derived once per system from the optk_surface
// records (host side, float64 composition of the transformations)
// ---------------------------------------------------------------------------

struct Affine32 {
  float m[9];
  float t[3];
  int32_t kind;
  int32_t pad;
};

struct FastSurface {
  optk_affine hop;             // previous surface frame -> this sag frame
  optk_affine sag_to_surface;  // this sag frame -> this surface frame
  optk_affine to_global;       // this surface frame -> system frame
  Affine32 aperture_from_sag;  // sag frame -> aperture frame (positions)
  Affine32 aperture_from_surface;  // surface frame -> aperture frame (directions)
  double c64, kp1_64, r64;     // curvature, 1 + conic, toroid radius of rotation
  float c, kp1, r, k;          // the same in float32 (k = conic constant)
  float surf_row[3], surf_tz;  // z row of sag_to_surface (toroid residual)
  int32_t sag_kind;
  int32_t polish;              // float64 Newton steps after the float32 solve
  // the aperture's parameters as its float32 test uses them: circle
  // {r^2}; sector {r^2, start, stop}; ellipse and rectangle {a, b} (r^2
  // rounded from its float64 square, each value rounded once)
  float ap32[4];
};

struct FastSpec {
  optk_affine first_hop[2];  // [0]: system frame -> sag frame of surface 0;
                             // [1]: the same after the object/system pre-maps
  optk_affine last[2];       // [0]: last surface frame -> system frame;
                             // [1]: -> the post frame (sensor)
  optk_affine first_skip;    // grid entry with surface 0 elided: input frame
                             // (before the pre-maps) -> sag frame of surface 1
  int32_t skip_first;        // 1: surface 0 is a no-op for grid rays
  int32_t pad;
};

#define OPTK_GRAZING_COS 0.25f   // multilayers: below this cosine, float64
#define OPTK_STEEP_SLOPE 0.5     // sag slope above which the intercept is polished

// host-side affine algebra (float64) --------------------------------------

struct HostAffine {
  double m[9];
  double t[3];
};

inline HostAffine haff_identity() {
  HostAffine a = {{1, 0, 0, 0, 1, 0, 0, 0, 1}, {0, 0, 0}};
  return a;
}

inline HostAffine haff_of(const optk_affine &a) {
  HostAffine h = haff_identity();
  if (a.kind == OPTK_XF_IDENTITY) return h;
  for (int k = 0; k < 3; ++k) h.t[k] = a.t[k];
  if (a.kind == OPTK_XF_AFFINE)
    for (int k = 0; k < 9; ++k) h.m[k] = a.m[k];
  return h;
}

// inverse of p -> m p + t for orthonormal m: p -> m^T p - m^T t
inline HostAffine haff_inverse(const HostAffine &a) {
  HostAffine h;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) h.m[3 * i + j] = a.m[3 * j + i];
  for (int i = 0; i < 3; ++i)
    h.t[i] = -(h.m[3 * i] * a.t[0] + h.m[3 * i + 1] * a.t[1] + h.m[3 * i + 2] * a.t[2]);
  return h;
}

// g after f: p -> g(f(p))
inline HostAffine haff_then(const HostAffine &f, const HostAffine &g) {
  HostAffine h;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j)
      h.m[3 * i + j] = g.m[3 * i] * f.m[j] + g.m[3 * i + 1] * f.m[3 + j] + g.m[3 * i + 2] * f.m[6 + j];
    h.t[i] = g.m[3 * i] * f.t[0] + g.m[3 * i + 1] * f.t[1] + g.m[3 * i + 2] * f.t[2] + g.t[i];
  }
  return h;
}

inline bool haff_is_identity_m(const HostAffine &a) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      if (a.m[3 * i + j] != (i == j ? 1.0 : 0.0)) return false;
  return true;
}

inline optk_affine haff_record(const HostAffine &a) {
  optk_affine r;
  for (int k = 0; k < 9; ++k) r.m[k] = a.m[k];
  for (int k = 0; k < 3; ++k) r.t[k] = a.t[k];
  r.pad = 0;
  bool no_t = a.t[0] == 0.0 && a.t[1] == 0.0 && a.t[2] == 0.0;
  if (haff_is_identity_m(a))
    r.kind = no_t ? OPTK_XF_IDENTITY : OPTK_XF_TRANSLATE;
  else
    r.kind = OPTK_XF_AFFINE;
  return r;
}

inline Affine32 haff_record32(const HostAffine &a) {
  Affine32 r;
  for (int k = 0; k < 9; ++k) r.m[k] = (float)a.m[k];
  for (int k = 0; k < 3; ++k) r.t[k] = (float)a.t[k];
  r.pad = 0;
  bool no_t = a.t[0] == 0.0 && a.t[1] == 0.0 && a.t[2] == 0.0;
  if (haff_is_identity_m(a))
    r.kind = no_t ? OPTK_XF_IDENTITY : OPTK_XF_TRANSLATE;
  else
    r.kind = OPTK_XF_AFFINE;
  return r;
}

// the pre-maps of the grid entry point: object transformation (forward),
// then the system transformation (inverse)
inline HostAffine haff_pre(const optk_affine pre[2], const int32_t pre_is_inverse[2]) {
  HostAffine h = haff_identity();
  for (int k = 0; k < 2; ++k) {
    if (pre[k].kind == OPTK_XF_IDENTITY) continue;
    HostAffine a = haff_of(pre[k]);
    h = haff_then(h, pre_is_inverse[k] ? haff_inverse(a) : a);
  }
  return h;
}

// largest |slope| of the sag over the region the aperture admits (a bound:
// the aperture's circumscribed radius about the sag axis); unbounded
// apertures count as steep
inline double sag_steepness(const optk_surface &s) {
  if (s.sag_kind == OPTK_SAG_NONE) return 0.0;
  double rad;
  switch (s.aperture_kind) {
  case OPTK_APERTURE_CIRCLE:
  case OPTK_APERTURE_SECTOR:
    rad = s.aperture_p[0];
    break;
  case OPTK_APERTURE_ELLIPSE:
  case OPTK_APERTURE_RECTANGLE:
    rad = sqrt(s.aperture_p[0] * s.aperture_p[0] + s.aperture_p[1] * s.aperture_p[1]);
    break;
  default:
    return INFINITY;  // polygons and open surfaces: no cheap bound
  }
  if (!(s.aperture_flags & OPTK_AP_ACTIVE) || (s.aperture_flags & OPTK_AP_INVERTED))
    return INFINITY;
  // offsets of the aperture and sag frames from the surface axis
  rad += sqrt(s.aperture_transform.t[0] * s.aperture_transform.t[0] +
              s.aperture_transform.t[1] * s.aperture_transform.t[1]);
  rad += sqrt(s.sag_transform.t[0] * s.sag_transform.t[0] +
              s.sag_transform.t[1] * s.sag_transform.t[1]);
  double c = fabs(1.0 / s.sag_p[0]);
  double cr = c * rad;
  switch (s.sag_kind) {
  case OPTK_SAG_CONIC: {
    double k = s.sag_p[1];
    double g = 1.0 - (1.0 + k) * cr * cr;
    return g > 0.0 ? cr / sqrt(g) : INFINITY;
  }
  case OPTK_SAG_PARABOLA:
    return cr;
  case OPTK_SAG_TOROID: {
    double cr2 = rad / fabs(s.sag_p[1]);
    double a = cr > cr2 ? cr : cr2;
    return a < 1.0 ? a / sqrt(1.0 - a * a) : INFINITY;
  }
  default: {  // sphere, cylinder
    return cr < 1.0 ? cr / sqrt(1.0 - cr * cr) : INFINITY;
  }
  }
}

// A surface that leaves a ray exactly as it is when the ray already lies on
// its (flat) sag plane: no sag, vacuum on both sides (index, efficiency and
// attenuation of the vacuum ray unchanged), no rulings, no active aperture,
// not a mirror.  Moving such a ray to the plane is t = -0 / u_z = -0, and a
// position plus -0 times a finite direction is the position itself.
inline bool fast_passthrough(const optk_surface &s) {
  return s.sag_kind == OPTK_SAG_NONE && s.material_kind == OPTK_MATERIAL_VACUUM &&
         s.rulings_kind == OPTK_RULINGS_NONE && !s.is_mirror &&
         (s.aperture_kind == OPTK_APERTURE_NONE || !(s.aperture_flags & OPTK_AP_ACTIVE));
}

// Build the float32-form records.  `surfaces` must already carry the derived
// constants (derive_constants).
inline void fast_build(const optk_surface *surfaces, int num_surfaces, const optk_affine pre[2],
                       const int32_t pre_is_inverse[2], const optk_affine &post,
                       FastSurface *out, FastSpec &spec) {
  spec.first_skip = haff_record(haff_identity());
  spec.skip_first = 0;
  spec.pad = 0;
  HostAffine prev_to_global = haff_identity();  // system frame to begin with
  for (int i = 0; i < num_surfaces; ++i) {
    const optk_surface &s = surfaces[i];
    FastSurface &f = out[i];
    memset(&f, 0, sizeof(f));
    HostAffine T = haff_of(s.transform);
    HostAffine S = haff_of(s.sag_transform);
    HostAffine A = haff_of(s.aperture_transform);
    // previous frame -> system -> surface -> sag
    HostAffine hop = haff_then(haff_then(prev_to_global, haff_inverse(T)), haff_inverse(S));
    f.hop = haff_record(hop);
    f.sag_to_surface = haff_record(S);
    f.to_global = haff_record(T);
    f.aperture_from_sag = haff_record32(haff_then(S, haff_inverse(A)));
    f.aperture_from_surface = haff_record32(haff_inverse(A));
    f.c64 = s.sag_c;
    f.kp1_64 = (s.sag_kind == OPTK_SAG_CONIC) ? 1.0 + s.sag_p[1]
                                              : (s.sag_kind == OPTK_SAG_PARABOLA ? 0.0 : 1.0);
    f.r64 = (s.sag_kind == OPTK_SAG_TOROID) ? s.sag_p[1] : s.sag_p[0];
    f.c = (float)f.c64;
    f.kp1 = (float)f.kp1_64;
    f.r = (float)f.r64;
    f.k = (float)(f.kp1_64 - 1.0);
    for (int k = 0; k < 3; ++k) f.surf_row[k] = (float)S.m[6 + k];
    f.surf_tz = (float)S.t[2];
    f.sag_kind = s.sag_kind;
    f.polish = sag_steepness(s) > OPTK_STEEP_SLOPE ? 2 : 0;
    if (s.aperture_kind == OPTK_APERTURE_CIRCLE || s.aperture_kind == OPTK_APERTURE_SECTOR) {
      f.ap32[0] = (float)(s.aperture_p[0] * s.aperture_p[0]);
      f.ap32[1] = (float)s.aperture_p[1];
      f.ap32[2] = (float)s.aperture_p[2];
    } else {
      f.ap32[0] = (float)s.aperture_p[0];
      f.ap32[1] = (float)s.aperture_p[1];
    }
    prev_to_global = T;
  }
  if (num_surfaces > 0) {
    HostAffine T0 = haff_of(surfaces[0].transform);
    HostAffine S0 = haff_of(surfaces[0].sag_transform);
    HostAffine to_sag0 = haff_then(haff_inverse(T0), haff_inverse(S0));
    spec.first_hop[0] = haff_record(to_sag0);
    spec.first_hop[1] = haff_record(haff_then(haff_pre(pre, pre_is_inverse), to_sag0));
    HostAffine Tl = haff_of(surfaces[num_surfaces - 1].transform);
    spec.last[0] = haff_record(Tl);
    spec.last[1] = haff_record(haff_then(Tl, haff_inverse(haff_of(post))));
    // Grid rays start on the plane z = 0 of the object frame.  When the
    // pre-maps and surface 0's frames carry that plane onto surface 0's sag
    // plane (the third row of the composed map is (0, 0, m33 | 0)) and
    // surface 0 is a pass-through, the grid kernels skip it: the input goes
    // straight to surface 1's sag frame.  The object surface of an imaging
    // system is such a surface.  (A ray parallel to the plane, u_z = 0, would
    // get t = -0/0 = NaN there; the kernels reproduce that case.)
    const HostAffine F1 = haff_then(haff_pre(pre, pre_is_inverse), to_sag0);
    if (num_surfaces > 1 && fast_passthrough(surfaces[0]) && F1.m[6] == 0.0 && F1.m[7] == 0.0 &&
        F1.t[2] == 0.0) {
      spec.first_skip = haff_record(haff_then(haff_then(F1, S0), haff_of(out[1].hop)));
      spec.skip_first = 1;
    }
  } else {
    HostAffine pre_h = haff_pre(pre, pre_is_inverse);
    spec.first_hop[0] = haff_record(haff_identity());
    spec.first_hop[1] = haff_record(pre_h);
    spec.last[0] = haff_record(haff_identity());
    spec.last[1] = haff_record(haff_inverse(haff_of(post)));
  }
}

// ---------------------------------------------------------------------------
// per-ray pieces
// ---------------------------------------------------------------------------

struct FRay {
  double px, py, pz;  // position (float64 across the hops)
  double dx, dy, dz;  // direction
  float wavelength, intensity, attenuation, index;
  bool unvignetted;
};

OPTK_HD void affine32_apply(const Affine32 &a, float &x, float &y, float &z) {
  if (a.kind == OPTK_XF_IDENTITY) return;
  if (a.kind == OPTK_XF_AFFINE) {
    float X = a.m[0] * x + a.m[1] * y + a.m[2] * z;
    float Y = a.m[3] * x + a.m[4] * y + a.m[5] * z;
    float Z = a.m[6] * x + a.m[7] * y + a.m[8] * z;
    x = X;
    y = Y;
    z = Z;
  }
  x += a.t[0];
  y += a.t[1];
  z += a.t[2];
}

// p -> m p + t and d -> m d under one record, in one branch on its kind,
// each row a chain of fused multiply-adds (the float32 form's transforms;
// the float64 form keeps xf_apply_pos / xf_apply_dir and numpy's rounding)
OPTK_HD void fast_xf(const optk_affine &a, double &px, double &py, double &pz, double &dx,
                     double &dy, double &dz) {
  if (a.kind == OPTK_XF_IDENTITY) return;
  if (a.kind == OPTK_XF_TRANSLATE) {
    px += a.t[0];
    py += a.t[1];
    pz += a.t[2];
    return;
  }
  const double *m = a.m;
  const double X = fma(m[0], px, fma(m[1], py, fma(m[2], pz, a.t[0])));
  const double Y = fma(m[3], px, fma(m[4], py, fma(m[5], pz, a.t[1])));
  const double Z = fma(m[6], px, fma(m[7], py, fma(m[8], pz, a.t[2])));
  const double DX = fma(m[0], dx, fma(m[1], dy, m[2] * dz));
  const double DY = fma(m[3], dx, fma(m[4], dy, m[5] * dz));
  const double DZ = fma(m[6], dx, fma(m[7], dy, m[8] * dz));
  px = X;
  py = Y;
  pz = Z;
  dx = DX;
  dy = DY;
  dz = DZ;
}

OPTK_HD void fast_xf(const optk_affine &a, FRay &r) { fast_xf(a, r.px, r.py, r.pz, r.dx, r.dy, r.dz); }

// Root of A s^2 + B s + C = 0 that the reference's closed forms select for
// spheres, cylinders and paraboloids: (-B - sign(u_z) sqrt(disc)) / (2A)
// (the sign(r u_z) / sign(f u_z) branch of _spherical.py:176-184,
// _cylindrical.py:147-151, _parabolic.py:142-151, rescaled by the curvature),
// evaluated without cancellation.
OPTK_HD float branch_root(float A, float B, float C, float uz) {
  float disc = B * B - 4.0f * A * C;
  float sq = sqrtf(disc);
  float sigma = uz > 0.0f ? 1.0f : -1.0f;
  float sb = B > 0.0f ? 1.0f : (B < 0.0f ? -1.0f : sigma);
  float q = -0.5f * (B + sb * sq);
  return (sb == sigma) ? q / A : C / q;
}

// float32 sag-residual solve from the vertex-plane point (X, Y, Z) along u;
// returns the path length s from that point (INFINITY when the conic has no
// admissible root, as the reference's t = inf).  t0 is the path length
// already travelled to reach the point (conic roots are chosen by |t0 + s|).
// the sag kind: compile-time (K, see Kinds) or the record's own
template <uint32_t K> OPTK_HD int k_fsag(const FastSurface &f) {
  return Kinds<K>::fixed ? Kinds<K>::sag : f.sag_kind;
}

template <uint32_t K = 0>
OPTK_HD float residual_intercept(const FastSurface &f, float X, float Y, float Z, float ux,
                                 float uy, float uz, double t0) {
  const float c = f.c;
  switch (k_fsag<K>(f)) {
  case OPTK_SAG_NONE:
    return -Z / uz;
  case OPTK_SAG_SPHERE: {
    float A = c * (ux * ux + uy * uy + uz * uz);
    float B = 2.0f * (c * (X * ux + Y * uy + Z * uz) - uz);
    float Cc = c * (X * X + Y * Y + Z * Z) - 2.0f * Z;
    return branch_root(A, B, Cc, uz);
  }
  case OPTK_SAG_CYLINDER: {
    float A = c * (ux * ux + uz * uz);
    float B = 2.0f * (c * (X * ux + Z * uz) - uz);
    float Cc = c * (X * X + Z * Z) - 2.0f * Z;
    float disc = B * B - 4.0f * A * Cc;
    return disc > 0.0f ? branch_root(A, B, Cc, uz) : -Z / uz;
  }
  case OPTK_SAG_PARABOLA: {
    float den = ux * ux + uy * uy;
    float A = c * den;
    float B = 2.0f * (c * (X * ux + Y * uy) - uz);
    float Cc = c * (X * X + Y * Y) - 2.0f * Z;
    return den > 1e-10f ? branch_root(A, B, Cc, uz) : -Cc / B;
  }
  case OPTK_SAG_CONIC: {
    const float kp1 = f.kp1;
    float A = c * (ux * ux + uy * uy + kp1 * (uz * uz));
    float B = 2.0f * (c * (X * ux + Y * uy + kp1 * Z * uz) - uz);
    float Cc = c * (X * X + Y * Y + kp1 * (Z * Z)) - 2.0f * Z;
    float disc = B * B - 4.0f * A * Cc;
    bool real = disc >= 0.0f;
    float s1, s2;
    if (fabsf(A) < 1e-12f) {  // degenerate: the linear root (_conic.py:140-145)
      s1 = s2 = -Cc / B;
    } else {
      float sq = sqrtf(real ? disc : 0.0f);
      float sb = B >= 0.0f ? 1.0f : -1.0f;
      float q = -0.5f * (B + sb * sq);
      s1 = q / A;
      s2 = Cc / q;
    }
    // vertex-sheet filter z (c r^2 - z) >= 0 (_conic.py:152-154), then the
    // root nearest the ray's origin (_conic.py:160-162)
    float z1 = Z + uz * s1, x1 = X + ux * s1, y1 = Y + uy * s1;
    float z2 = Z + uz * s2, x2 = X + ux * s2, y2 = Y + uy * s2;
    bool ok1 = real && z1 * (c * (x1 * x1 + y1 * y1) - z1) >= 0.0f;
    bool ok2 = real && z2 * (c * (x2 * x2 + y2 * y2) - z2) >= 0.0f;
    if (!ok1 && !ok2) return INFINITY;
    double t1 = ok1 ? t0 + (double)s1 : INFINITY;
    double t2 = ok2 ? t0 + (double)s2 : INFINITY;
    return fabs(t1) <= fabs(t2) ? s1 : s2;
  }
  }
  return -Z / uz;
}

// compensated toroid sag: r - sqrt((r - zy)^2 - x^2) written as
// (zy (2r - zy) + x^2) / (r + sqrt(...)) for r > 0 (no cancellation)
template <typename T> struct ToroidEval {
  T z, dzdx, dzdy;
};

// fast square root and reciprocal per real type
template <typename T> struct FastOps;
template <> struct FastOps<float> {
  static OPTK_HD float sqrt(float x) { return sqrtf(x); }
  static OPTK_HD float rcp(float x) { return 1.0f / x; }
  static OPTK_HD float rsqrt(float x) { return 1.0f / sqrtf(x); }
};
template <> struct FastOps<double> {
  static OPTK_HD double sqrt(double x) { return sqrt64(x); }
  static OPTK_HD double rcp(double x) { return rcp64(x); }
  // arguments of the normals' reciprocal square roots are >= 1 (toroid) or
  // positive wherever the conic's sag is real (1 - k c^2 r^2 > 0 when
  // (1 + k) c^2 r^2 <= 1); elsewhere the normal is NaN either way
  static OPTK_HD double rsqrt(double x) { return rsqrt64_finite(x); }
};

template <typename T> OPTK_HD ToroidEval<T> toroid_eval(T c, T r, T x, T y) {
  using Op = FastOps<T>;
  T y2 = y * y;
  T g = Op::sqrt(T(1) - c * c * y2);
  T zy = c * y2 * Op::rcp(T(1) + g);
  T rz = r - zy;
  T f = Op::sqrt(rz * rz - x * x);
  T z = r > T(0) ? (zy * (T(2) * r - zy) + x * x) * Op::rcp(r + f) : r - f;
  T inv_f = Op::rcp(f);
  return {z, x * inv_f, rz * (c * y * Op::rcp(g)) * inv_f};
}

// toroid: Newton on g(s) = z_surface(s) - sag(x(s), y(s)) (the residual of
// AbstractSag.intercept, optika/sags/_abc.py:76-107, whose z is taken in the
// surface frame and whose sag is evaluated in the sag frame)
OPTK_HD float toroid_residual_solve(const FastSurface &f, float X, float Y, float Z, float ux,
                                    float uy, float uz) {
  const float dzs = f.surf_row[0] * ux + f.surf_row[1] * uy + f.surf_row[2] * uz;
  float s = 0.0f;
  for (int it = 0; it < 24; ++it) {
    float x = X + ux * s, y = Y + uy * s, z = Z + uz * s;
    ToroidEval<float> e = toroid_eval<float>(f.c, f.r, x, y);
    float zs = f.surf_row[0] * x + f.surf_row[1] * y + f.surf_row[2] * z + f.surf_tz;
    float step = (zs - e.z) / (dzs - (e.dzdx * ux + e.dzdy * uy));
    s -= step;
    if (fabsf(step) <= 1e-6f) break;
  }
  return s;
}

// two float64 Newton steps on the residual of a steep sag, from the float32
// solution s (steep conics and toroids)
template <uint32_t K = 0>
OPTK_HD double polish64(const FastSurface &f, double X, double Y, double Z, double ux, double uy,
                        double uz, double s) {
  for (int it = 0; it < 2; ++it) {
    double x = X + ux * s, y = Y + uy * s, z = Z + uz * s;
    double g, dg;
    if (k_fsag<K>(f) == OPTK_SAG_TOROID) {
      ToroidEval<double> e = toroid_eval<double>(f.c64, f.r64, x, y);
      double zs = (double)f.surf_row[0] * x + (double)f.surf_row[1] * y +
                  (double)f.surf_row[2] * z + (double)f.surf_tz;
      g = zs - e.z;
      dg = (double)f.surf_row[0] * ux + (double)f.surf_row[1] * uy + (double)f.surf_row[2] * uz -
           (e.dzdx * ux + e.dzdy * uy);
    } else {
      // implicit conic c (x^2 + y^2 + (1+k) z^2) - 2 z (cylinder: no y)
      double yy = k_fsag<K>(f) == OPTK_SAG_CYLINDER ? 0.0 : y;
      double uyy = k_fsag<K>(f) == OPTK_SAG_CYLINDER ? 0.0 : uy;
      g = f.c64 * (x * x + yy * yy + f.kp1_64 * z * z) - 2.0 * z;
      dg = 2.0 * (f.c64 * (x * ux + yy * uyy + f.kp1_64 * z * uz) - uz);
    }
    s = s - g * rcp64(dg);
  }
  return s;
}

// path length from the current (sag-frame) position to the sag
template <uint32_t K = 0>
OPTK_HD double fast_intercept(const FastSurface &f, const FRay &r) {
  // hop to the vertex plane z = 0 in float64 (a ray parallel to it stays put)
  double t0 = fabs(r.dz) > 1e-6 ? -r.pz * rcp64_finite(r.dz) : 0.0;
  if (k_fsag<K>(f) == OPTK_SAG_NONE && fabs(r.dz) > 1e-6) return t0;
  double X = r.px + r.dx * t0, Y = r.py + r.dy * t0, Z = r.pz + r.dz * t0;
  float ux = (float)r.dx, uy = (float)r.dy, uz = (float)r.dz;
  float s;
  if (k_fsag<K>(f) == OPTK_SAG_TOROID)
    s = toroid_residual_solve(f, (float)X, (float)Y, (float)Z, ux, uy, uz);
  else
    s = residual_intercept<K>(f, (float)X, (float)Y, (float)Z, ux, uy, uz, t0);
  double sd = (double)s;
  if (f.polish && k_fsag<K>(f) != OPTK_SAG_NONE && fabs(sd) < INFINITY)
    sd = polish64<K>(f, X, Y, Z, r.dx, r.dy, r.dz, sd);
  return t0 + sd;
}

// Normal from sag-frame coordinates (optika/sags/*.py normal()), with the
// fast square roots of FastOps<T>.  The trace evaluates it in float64 at the
// float64 hit point: a float32 normal tilts the outgoing ray by ~1e-8 rad,
// which over a metre-long arm is ~1e-3 pixel.
template <typename T, uint32_t K = 0>
OPTK_HD V3<T> fast_normal(const FastSurface &f, T x, T y) {
  using Op = FastOps<T>;
  const T c = sizeof(T) == 8 ? T(f.c64) : T(f.c);
  const T kp1 = sizeof(T) == 8 ? T(f.kp1_64) : T(f.kp1);
  switch (k_fsag<K>(f)) {
  case OPTK_SAG_SPHERE: {
    T nx = c * x, ny = c * y;
    return {nx, ny, -Op::sqrt(T(1) - nx * nx - ny * ny)};
  }
  case OPTK_SAG_CONIC:
  case OPTK_SAG_PARABOLA: {
    // (c x / g, c y / g, -1) / |.| = (c x, c y, -g) / sqrt(1 - k c^2 r^2);
    // a paraboloid (1 + k = 0, uniform per surface) has g = 1 exactly
    T c2r2 = c * c * (x * x + y * y);
    T g = kp1 == T(0) ? T(1) : Op::sqrt(T(1) - kp1 * c2r2);
    T inv = Op::rsqrt(T(1) - (kp1 - T(1)) * c2r2);
    return {c * x * inv, c * y * inv, -g * inv};
  }
  case OPTK_SAG_CYLINDER: {
    T nx = c * x;
    return {nx, T(0), -Op::sqrt(T(1) - nx * nx)};
  }
  case OPTK_SAG_TOROID: {
    const T r = sizeof(T) == 8 ? T(f.r64) : T(f.r);
    ToroidEval<T> e = toroid_eval<T>(c, r, x, y);
    T inv = Op::rsqrt(e.dzdx * e.dzdx + e.dzdy * e.dzdy + T(1));
    return {e.dzdx * inv, e.dzdy * inv, -inv};
  }
  }
  return {T(0), T(0), T(-1)};
}

// Snell's law (_snells_law.py:341-366) in float64 with fast reciprocal and
// square root
// `unit`: |a| = 1 (no grating before): with an unchanged index the law
// reduces exactly to b = a (refraction) or b = a - 2 (a.u) u (mirror), the
// square root of 1 + (a.u)^2 - |a|^2 = |a.u| being taken algebraically.
OPTK_HD V3<double> fast_snell(V3<double> a, double n1, double n2, V3<double> u, bool mirror,
                              bool unit) {
  bool same = n1 == n2 && n1 != 0.0 && fabs(n1) < INFINITY;
  if (same && unit) {
    if (!mirror) return a;
    const double au2 = 2.0 * (a.x * u.x + a.y * u.y + a.z * u.z);
    return {a.x - au2 * u.x, a.y - au2 * u.y, a.z - au2 * u.z};
  }
  double a2 = a.x * a.x + a.y * a.y + a.z * a.z;
  double r = same ? 1.0 : n1 * rcp64(n2);
  double au = a.x * u.x + a.y * u.y + a.z * u.z;
  double sgn = -copysign(1.0, au);
  double inv_r2 = same ? 1.0 : rcp64(r * r);
  double d = -au + sgn * (mirror ? 1.0 : -1.0) * sqrt64(inv_r2 + au * au - a2);
  return {r * (a.x + d * u.x), r * (a.y + d * u.y), r * (a.z + d * u.z)};
}

// Sellmeier glass (_materials.py:428-438), fast float64
template <uint32_t K = 0>
OPTK_HD double fast_index(const optk_surface &s, double wavelength, double n1) {
  switch (k_material<K>(s)) {
  case OPTK_MATERIAL_VACUUM:
    return 1.0;
  case OPTK_MATERIAL_GLASS: {
    double w2 = wavelength * wavelength;
    const double *p = s.material_p;
    double n2 = 1.0 + (p[0] * w2 * rcp64(w2 - p[3]) + p[1] * w2 * rcp64(w2 - p[4]) +
                       p[2] * w2 * rcp64(w2 - p[5]));
    return sqrt64(n2);
  }
  default:
    return n1;
  }
}

// ---------------------------------------------------------------------------
// float32 refractive index tables
// ---------------------------------------------------------------------------

OPTK_HD Cf chemical_nf(const optk_chemical &ch, const float *tables, float wavelength_nm,
                       const ChemGrid *grid, const int32_t *lo_tab) {
  const float *w = tables + ch.offset;
  const float *n = w + ch.count;
  const float *k = n + ch.count;
  const int cnt = ch.count;
  float x = wavelength_nm * 10.0f;
  if (!(x >= w[0])) return (x != x) ? Cf{x, x} : Cf{n[0], k[0]};
  if (x >= w[cnt - 1]) return {n[cnt - 1], k[cnt - 1]};
  int lo;
  if (grid && grid->cells) {
    lo = grid_bracket(*grid, lo_tab, w, cnt, x);
  } else {
    lo = 0;
    int hi = cnt - 1;
    while (hi - lo > 1) {
      int mid = (lo + hi) >> 1;
      if (x < w[mid])
        hi = mid;
      else
        lo = mid;
    }
  }
  float inv_dx = 1.0f / (w[lo + 1] - w[lo]);
  float dx = x - w[lo];
  return {(n[lo + 1] - n[lo]) * inv_dx * dx + n[lo], (k[lo + 1] - k[lo]) * inv_dx * dx + k[lo]};
}

// ---------------------------------------------------------------------------
// float32 multilayer: both polarizations in one pass
// ---------------------------------------------------------------------------

struct MLMedium {
  Cf n, cos;     // index and cosine (Snell invariant n sin)
  Cf qs, qp;     // s: n cos ; p: conj(cos) / n
  Cf u00, u11;   // propagation phases exp(-i beta), exp(+i beta)
  bool ok;       // |u00| < 1e10 (the reference's validity guard)
};

struct MLPair {
  M2<float> s, p;
};

OPTK_HD MLMedium ml_medium(Cf n, Cf kappa2, float thickness, float wavelength) {
  MLMedium m;
  m.n = n;
  Cf ratio = fcdiv(kappa2, cmul(n, n));
  m.cos = fcsqrt(Cf{1.0f - ratio.re, -ratio.im});
  m.qs = cmul(m.cos, n);
  m.qp = fcdiv(cconj(m.cos), n);
  if (thickness == 0.0f) {
    m.u00 = {1.0f, 0.0f};
    m.u11 = {1.0f, 0.0f};
    m.ok = true;
    return m;
  }
  // beta = 2 pi h n cos / wavelength (matrices.py:241-246)
  Cf nc = m.qs;
  float scale = 6.283185307179586f * thickness / wavelength;
  float bre = scale * nc.re, bim = scale * nc.im;
  float sb, cb;
  sincosf(bre, &sb, &cb);
  float e = expf(bim);
  float ei = expf(-bim);
  m.u00 = {e * cb, -e * sb};
  m.u11 = {ei * cb, ei * sb};
  m.ok = e < 1e10f;
  return m;
}

OPTK_HD float ml_profile(int kind, float width, float wavelength, Cf di, Cf n) {
  if (kind == OPTK_PROFILE_NONE) return 1.0f;
  // s = Re(4 pi n di / wavelength) (profiles.py:103-126)
  float s = 12.566370614359172f * (n.re * di.re - n.im * di.im) / wavelength;
  switch (kind) {
  case OPTK_PROFILE_ERF: {
    float sw = s * width;
    return expf(-(sw * sw) * 0.5f);
  }
  case OPTK_PROFILE_EXPONENTIAL: {
    float sw = s * width;
    return 1.0f / (1.0f + (sw * sw) * 0.5f);
  }
  case OPTK_PROFILE_LINEAR: {
    float x = 1.7320508075688772f * width * s;
    return sinf(x) / x;
  }
  case OPTK_PROFILE_SINUSOIDAL: {
    const float pi = 3.14159265358979f;
    float a = pi / (pi * pi - 8.0f);
    float x = a * width * s;
    float x1 = x - pi / 2.0f, x2 = x + pi / 2.0f;
    return pi * (sinf(x1) / x1 + sinf(x2) / x2) / 4.0f;
  }
  }
  return 1.0f;
}

// interface i -> j followed by propagation through j:
// (1/t) [[1, r], [r, 1]] diag(u00, u11) with 1/t = (qi + qj) / (2 qi) and
// r/t = f (qi - qj) / (2 qi)
OPTK_HD M2<float> ml_step(Cf qi, Cf qj, float f, Cf u00, Cf u11) {
  Cf w = fcrcp(cscale(2.0f, qi));
  Cf A = cmul(cadd(qi, qj), w);
  Cf B = cscale(f, cmul(csub(qi, qj), w));
  return {cmul(A, u00), cmul(B, u11), cmul(B, u00), cmul(A, u11)};
}

OPTK_HD MLPair ml_transfer(const MLMedium &i, const MLMedium &j, const optk_layer &layer_j,
                           float wavelength) {
  float fs = 1.0f, fp = 1.0f;
  if (layer_j.profile_kind != OPTK_PROFILE_NONE) {
    fs = ml_profile(layer_j.profile_kind, (float)layer_j.profile_width, wavelength, i.cos, i.n);
    fp = ml_profile(layer_j.profile_kind, (float)layer_j.profile_width, wavelength, cconj(i.cos),
                    i.n);
  }
  return {ml_step(i.qs, j.qs, fs, j.u00, j.u11), ml_step(i.qp, j.qp, fp, j.u00, j.u11)};
}

OPTK_HD MLPair ml_mul(const MLPair &x, const MLPair &y) { return {mmul(x.s, y.s), mmul(x.p, y.p)}; }

OPTK_HD MLPair ml_identity() { return {midentity<float>(), midentity<float>()}; }

OPTK_HD MLPair ml_power(MLPair m, int exponent) {
  MLPair result = ml_identity();
  bool have = false;
  while (exponent > 0) {
    if (exponent & 1) {
      result = have ? ml_mul(result, m) : m;
      have = true;
    }
    exponent >>= 1;
    if (exponent) m = ml_mul(m, m);
  }
  return result;
}

OPTK_HD Cf ml_layer_n(const optk_layer &L, const StackView &sv, const float *tables32,
                      float wavelength) {
  if (L.chemical < 0) return {1.0f, 0.0f};
  return chemical_nf(sv.chemicals[L.chemical], tables32, wavelength,
                     sv.grids ? sv.grids + L.chemical : nullptr, sv.grid_lo);
}

// Reflectance (mirror) or transmittance (film) averaged over s and p, or a
// negative value when the float32 evaluation does not apply (grazing
// incidence, the validity guard, more than 4 layers in a periodic block).
// im_q_sub (optional, film only): Im(n_sub cos_sub) of the wave refracted
// into the substrate
OPTK_HD float fast_multilayer(const StackView &sv, const float *tables32, int stack_index,
                              float wavelength, Cf cos_amb, Cf n_amb, bool film,
                              float *im_q_sub = nullptr) {
  if (fabsf(cos_amb.re) < OPTK_GRAZING_COS) return -1.0f;
  const optk_stack &stk = sv.stacks[stack_index];
  // Snell invariant: kappa^2 = n_amb^2 (1 - cos_amb^2)
  Cf one_minus = {1.0f - (cos_amb.re * cos_amb.re - cos_amb.im * cos_amb.im),
                  -2.0f * cos_amb.re * cos_amb.im};
  Cf kappa2 = cmul(cmul(n_amb, n_amb), one_minus);
  MLMedium amb;
  amb.n = n_amb;
  amb.cos = cos_amb;
  amb.qs = cmul(cos_amb, n_amb);
  amb.qp = fcdiv(cconj(cos_amb), n_amb);
  amb.ok = true;
  MLMedium prev = amb;
  MLPair m = ml_identity();
  bool ok = true;
  for (int b = 0; b < stk.num_blocks; ++b) {
    const optk_block &blk = stk.blocks[b];
    MLPair x;
    if (blk.kind == OPTK_BLOCK_PERIODIC) {
      if (blk.num_layers < 1 || blk.num_layers > 4) return -1.0f;
      const optk_layer &L1 = sv.layers[blk.first_layer];
      MLMedium first = ml_medium(ml_layer_n(L1, sv, tables32, wavelength), kappa2,
                                 (float)L1.thickness, wavelength);
      ok = ok && first.ok;
      // layers 2..m, shared by the first period and the repeated period
      MLPair inner = ml_identity();
      MLMedium last = first;
      for (int j = 1; j < blk.num_layers; ++j) {
        const optk_layer &L = sv.layers[blk.first_layer + j];
        MLMedium med = ml_medium(ml_layer_n(L, sv, tables32, wavelength), kappa2,
                                 (float)L.thickness, wavelength);
        ok = ok && med.ok;
        MLPair st = ml_transfer(last, med, L, wavelength);
        inner = (j == 1) ? st : ml_mul(inner, st);
        last = med;
      }
      MLPair start = ml_transfer(prev, first, L1, wavelength);
      MLPair period = ml_transfer(last, first, L1, wavelength);
      if (blk.num_layers > 1) {
        start = ml_mul(start, inner);
        period = ml_mul(period, inner);
      }
      x = (blk.num_periods > 1) ? ml_mul(start, ml_power(period, blk.num_periods - 1)) : start;
      prev = last;
    } else {
      x = ml_identity();
      for (int j = 0; j < blk.num_layers; ++j) {
        const optk_layer &L = sv.layers[blk.first_layer + j];
        MLMedium med = ml_medium(ml_layer_n(L, sv, tables32, wavelength), kappa2,
                                 (float)L.thickness, wavelength);
        ok = ok && med.ok;
        x = ml_mul(x, ml_transfer(prev, med, L, wavelength));
        prev = med;
      }
    }
    m = (b == 0) ? x : ml_mul(m, x);
  }
  if (!ok) return -1.0f;
  // substrate: interface only (thickness zero)
  optk_layer sub = {-1, OPTK_PROFILE_NONE, 0.0, 0.0};
  if (stk.substrate_layer >= 0) sub = sv.layers[stk.substrate_layer];
  MLMedium msub = ml_medium(ml_layer_n(sub, sv, tables32, wavelength), kappa2, 0.0f, wavelength);
  m = ml_mul(m, ml_transfer(prev, msub, sub, wavelength));
  if (!film) {
    float rs = fabs2(m.s.c) / fabs2(m.s.a);
    float rp = fabs2(m.p.c) / fabs2(m.p.a);
    return 0.5f * (rs + rp);
  }
  float ts = 1.0f / fabs2(m.s.a);
  float tp = 1.0f / fabs2(m.p.a);
  float gs = fcdiv(msub.qs, amb.qs).re;
  float gp = fcdiv(msub.qp, amb.qp).re;
  if (im_q_sub) *im_q_sub = msub.qs.im;
  return 0.5f * (ts * gs + tp * gp);
}

// The float64 multilayer (optk_physics.h) for the rays the float32 pass
// hands back.  Kept out of line on the device: its registers are then
// allocated for this rare path alone instead of raising the pressure (and
// the spills) of every ray's main path.
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __attribute__((noinline))
#else
inline
#endif
float multilayer_fallback64(const StackView &sv, int stack, float wavelength, double cos_theta,
                            float index, float attenuation, bool film) {
  double k64 = (double)attenuation * (double)wavelength / (4.0 * M_PI) * 1e-6;
  MultilayerRT<double> rt = multilayer_efficiency<double>(
      sv, stack, (double)wavelength, C<double>{cos_theta, 0.0}, C<double>{(double)index, k64}, film);
  return (float)(film ? (rt.T[0] + rt.T[1]) / 2.0 : (rt.R[0] + rt.R[1]) / 2.0);
}

// the silicon sensor's absorbance in float64, for the same rays
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __attribute__((noinline))
#else
inline
#endif
float silicon_fallback64(const StackView &sv, int stack, double thickness, float wavelength,
                         double cos_theta, float index, float attenuation) {
  double k64 = (double)attenuation * (double)wavelength / (4.0 * M_PI) * 1e-6;
  return (float)silicon_absorbance_average(sv, stack, thickness, (double)wavelength,
                                           C<double>{cos_theta, 0.0},
                                           C<double>{(double)index, k64});
}

template <int F, uint32_t K = 0>
OPTK_HD float fast_material_efficiency(const optk_surface &s, const StackView &sv,
                                      const float *tables32, float wavelength, V3<double> a,
                                      float index, float attenuation, V3<double> normal) {
  switch (k_material<K>(s)) {
  case OPTK_MATERIAL_MEASURED_MIRROR: {
    const float *w = tables32 + s.material_table;
    return interp1f(w, w + s.material_table_len, s.material_table_len, wavelength);
  }
  case OPTK_MATERIAL_MULTILAYER_MIRROR:
  case OPTK_MATERIAL_MULTILAYER_FILM: {  // _multilayers.py:839-935
    if (!(F & FEAT_MULTILAYER)) return 1.0f;
    float k = attenuation * wavelength / (4.0f * 3.14159265358979f) * 1e-6f;
    float cosv = (float)(-(a.x * normal.x + a.y * normal.y + a.z * normal.z));
    bool film = k_material<K>(s) == OPTK_MATERIAL_MULTILAYER_FILM;
    float eff = fast_multilayer(sv, tables32, s.stack, wavelength, Cf{cosv, 0.0f},
                                Cf{index, k}, film);
    if (eff < 0.0f) {
      // grazing incidence or the validity guard: the float64 routine
      double c64 = -(a.x * normal.x + a.y * normal.y + a.z * normal.z);
      eff = multilayer_fallback64(sv, s.stack, wavelength, c64, index, attenuation, film);
    }
    return eff;
  }
  case OPTK_MATERIAL_SILICON_SENSOR: {  // _materials.py:2231-2243
    if (!(F & FEAT_MULTILAYER)) return 1.0f;
    float k = attenuation * wavelength / (4.0f * 3.14159265358979f) * 1e-6f;
    float cosv = (float)(-(a.x * normal.x + a.y * normal.y + a.z * normal.z));
    float im_q = 0.0f;
    float t = fast_multilayer(sv, tables32, s.stack, wavelength, Cf{cosv, 0.0f}, Cf{index, k},
                              true, &im_q);
    if (t < 0.0f) {
      double c64 = -(a.x * normal.x + a.y * normal.y + a.z * normal.z);
      return silicon_fallback64(sv, s.stack, s.material_p[0], wavelength, c64, index, attenuation);
    }
    // the decay 1 - exp(-alpha d) is the same for both polarizations
    float absorption = 4.0f * 3.14159265358979f * im_q / wavelength;
    return t - t * expf(-absorption * (float)s.material_p[0]);
  }
  default:
    return 1.0f;
  }
}

template <uint32_t K = 0>
OPTK_HD bool fast_aperture(const optk_surface &s, const float (&ap)[4], const Affine32 &to_aperture,
                           const float *tables32, float x, float y, float z) {
  affine32_apply(to_aperture, x, y, z);
  bool mask = true;
  switch (k_aperture<K>(s)) {
  case OPTK_APERTURE_CIRCLE:
    mask = x * x + y * y <= ap[0];
    break;
  case OPTK_APERTURE_SECTOR: {
    bool mr = x * x + y * y <= ap[0];
    float ang = atan2f(y, x);
    const float two_pi = 6.283185307179586f;
    float apos = fmodf(ang, two_pi);
    if (apos < 0.0f) apos += two_pi;
    float aneg = fmodf(ang, -two_pi);
    if (aneg > 0.0f) aneg -= two_pi;
    float a0 = ap[1], a1 = ap[2];
    mask = mr && (((a0 < apos) && (apos < a1)) || ((a0 < aneg) && (aneg < a1)));
    break;
  }
  case OPTK_APERTURE_ELLIPSE: {
    float u = x / ap[0], v = y / ap[1];
    mask = u * u + v * v <= 1.0f;
    break;
  }
  case OPTK_APERTURE_RECTANGLE: {
    float hx = ap[0], hy = ap[1];
    mask = ((-hx <= x) && (x <= hx)) && ((-hy <= y) && (y <= hy));
    break;
  }
  case OPTK_APERTURE_POLYGON: {
    const float *vx = tables32 + s.polygon_offset;
    const float *vy = vx + s.polygon_count;
    int nv = s.polygon_count;
    bool inside = false;
    int j = nv - 1;
    for (int i = 0; i < nv; ++i) {
      float xi = vx[i], yi = vy[i], xj = vx[j], yj = vy[j];
      bool cond = (yi > y) != (yj > y);
      float xc = (xj - xi) * (y - yi) / (yj - yi) + xi;
      if (cond && (x < xc)) inside = !inside;
      j = i;
    }
    mask = inside;
    break;
  }
  default:
    return true;
  }
  const int flags = k_aperture_flags<K>(s);
  if (flags & OPTK_AP_INVERTED) mask = !mask;
  if (!(flags & OPTK_AP_ACTIVE)) mask = true;
  return mask;
}

struct FastView {
  const optk_surface *surfaces;
  const FastSurface *fast;
  StackView sv;
  const float *tables32;
};

// one surface of the float32 form: the ray arrives in this surface's sag
// frame (the caller applies the hop) and leaves in its surface frame
template <int F, uint32_t K = 0>
OPTK_HD void fast_propagate(const FastSurface &f, const optk_surface &s, const FastView &v,
                            int path, FRay &r) {
  double t = fast_intercept<K>(f, r);
  double hx = r.px + r.dx * t, hy = r.py + r.dy * t, hz = r.pz + r.dz * t;
  // attenuation over the path (|d| = 1); exactly 1 for alpha = 0
  if (!(r.attenuation == 0.0f && fabs(t) < 1e150)) {
    float len = (float)fabs(t) * sqrtf((float)(r.dx * r.dx + r.dy * r.dy + r.dz * r.dz));
    r.intensity = expf(-r.attenuation * len) * r.intensity;
  }
  V3<double> normal = fast_normal<double, K>(f, hx, hy);
  float sx = (float)hx, sy = (float)hy, sz = (float)hz;  // sag frame, float

  // back to the surface frame
  fast_xf(f.sag_to_surface, hx, hy, hz, r.dx, r.dy, r.dz);
  r.px = hx;
  r.py = hy;
  r.pz = hz;

  V3<double> a = {r.dx, r.dy, r.dz};
  int order = 0;
  const bool ruled = (F & FEAT_RULINGS) && k_rulings<K>(s) != OPTK_RULINGS_NONE;
  if (ruled) {
    order = path_order(s, path);
    V3<double> p1 = {hx, hy, hz};
    V3<double> kappa = spacing_vector<double, K>(s, p1, normal);
    a = incident_effective(s, order, (double)r.wavelength, a, (double)r.index, normal, kappa);
  }

  const double n1 = r.index;
  const double n2 = fast_index<K>(s, (double)r.wavelength, n1);
  V3<double> b = fast_snell(a, n1, n2, normal, k_mirror<K>(s), !ruled);
  float eff = fast_material_efficiency<F, K>(s, v.sv, v.tables32, r.wavelength, a, r.index,
                                             r.attenuation, normal);
  if (ruled) {
    V3<float> p1 = {(float)hx, (float)hy, (float)hz};
    V3<float> af = {(float)a.x, (float)a.y, (float)a.z};
    V3<float> nf = {(float)normal.x, (float)normal.y, (float)normal.z};
    eff = eff * rulings_efficiency<float, K>(s, order, r.wavelength, af, p1, nf, v.tables32);
  }
  r.intensity = r.intensity * eff;
  r.attenuation = (float)material_attenuation<K>(s, (double)r.attenuation);
  if (!(n1 == n2 && n1 != 0.0 && fabs(n1) < INFINITY))
    r.wavelength = (float)((double)r.wavelength * n2 * rcp64(n1));
  r.index = (float)n2;
  r.dx = b.x;
  r.dy = b.y;
  r.dz = b.z;

  if (k_aperture<K>(s) != OPTK_APERTURE_NONE) {
    bool m = (k_aperture_flags<K>(s) & OPTK_AP_DIRECTION)
                 ? fast_aperture<K>(s, f.ap32, f.aperture_from_surface, v.tables32, (float)r.dx,
                                    (float)r.dy, (float)r.dz)
                 : fast_aperture<K>(s, f.ap32, f.aperture_from_sag, v.tables32, sx, sy, sz);
    r.unvignetted = r.unvignetted && m;
  }
}

// sin and cos of a float64 angle by Maclaurin series in Horner form, the
// reciprocal factorial ratios as constants (no division instructions):
// through x^5 / x^6 for |x| <= 0.01 rad (next terms x^7/5040 and x^8/40320:
// below 2.5e-16 relative), through x^11 / x^12 for |x| <= 0.2 rad
// (truncation below 1e-18 relative); beyond, the float32 library functions
// (about 1e-7 relative: field angles of imaging systems are small, and the
// float64 library sincos would cost the whole kernel registers for a branch
// it rarely takes).  A wave whose lanes all take the first branch skips the
// others.
OPTK_HD void small_sincos(double x, double &s, double &c) {
  const double x2 = x * x;
  if (fabs(x) <= 0.01) {
    s = x * (1.0 - x2 * (1.0 / 6.0) * (1.0 - x2 * (1.0 / 20.0)));
    c = 1.0 - x2 * 0.5 * (1.0 - x2 * (1.0 / 12.0) * (1.0 - x2 * (1.0 / 30.0)));
    return;
  }
  if (fabs(x) <= 0.2) {
    s = x * (1.0 - x2 * (1.0 / 6.0) * (1.0 - x2 * (1.0 / 20.0) * (1.0 - x2 * (1.0 / 42.0) *
        (1.0 - x2 * (1.0 / 72.0) * (1.0 - x2 * (1.0 / 110.0))))));
    c = 1.0 - x2 * 0.5 * (1.0 - x2 * (1.0 / 12.0) * (1.0 - x2 * (1.0 / 30.0) * (1.0 - x2 * (1.0 / 56.0) *
        (1.0 - x2 * (1.0 / 90.0) * (1.0 - x2 * (1.0 / 132.0))))));
    return;
  }
  float sf, cf;
  sincosf((float)x, &sf, &cf);
  s = sf;
  c = cf;
}

// optika.direction (optika/_util.py:41-73) for the float32 inputs, in
// float64
OPTK_HD void fast_direction(float ax_deg, float ay_deg, double &dx, double &dy, double &dz) {
  const double deg = 0.017453292519943295;  // pi / 180
  double sx, cx, sy, cy;
  small_sincos((double)ax_deg * deg, sx, cx);
  small_sincos((double)ay_deg * deg, sy, cy);
  dx = -cy * sx;
  dy = -sy;
  dz = cy * cx;
}

OPTK_HD FRay fast_grid_ray(float fx, float fy, float ux, float uy, float wavelength,
                           float intensity, int object_at_infinity) {
  FRay r;
  if (object_at_infinity) {
    r.px = ux;
    r.py = uy;
    fast_direction(fx, fy, r.dx, r.dy, r.dz);
  } else {
    r.px = fx;
    r.py = fy;
    fast_direction(ux, uy, r.dx, r.dy, r.dz);
  }
  r.pz = 0.0;
  r.wavelength = wavelength;
  r.intensity = intensity;
  r.attenuation = 0.0f;
  r.index = 1.0f;
  r.unvignetted = true;
  return r;
}

}  // namespace optk
