// optk_surface.h -- one surface step of the sequential trace (float64):
// AbstractSurface.propagate_rays (optika/surfaces.py:123-198) with every
// sag, ruling, material and aperture operator inlined.
#pragma once

#include "optk_physics.h"

namespace optk {

// Compile-time feature masks of the trace kernels: a system that needs no
// toroid Newton solve, no rulings or no multilayer transfer matrices runs a
// kernel without that code, which keeps its register budget (and hence its
// occupancy) at what the remaining operators need.
enum : int {
  FEAT_TOROID = 1,
  FEAT_RULINGS = 2,
  FEAT_MULTILAYER = 4,
  FEAT_ALL = 7,
};

// Compile-time kinds of one surface record.  A specialised kernel fixes the
// kinds of every surface of a system at compile time (K != 0, built by
// kinds_pack): each per-surface switch then folds to the one branch the
// surface takes, and what the fixed kinds make constant -- an index of 1
// through vacuum and mirrors, a zero attenuation, a unit index ratio --
// propagates through the whole trace.  K == 0 reads every kind from the
// record, as the generic kernels do.  The arithmetic of each branch is the
// same either way.
constexpr uint32_t KINDS_FIXED = 0x80000000u;

constexpr uint32_t kinds_pack(int xf, int sag, int sag_xf, int material, int mirror, int rulings,
                              int spacing, int aperture, int aperture_xf, int aperture_flags) {
  return KINDS_FIXED | (uint32_t)xf | ((uint32_t)sag << 2) | ((uint32_t)sag_xf << 5) |
         ((uint32_t)material << 7) | ((uint32_t)mirror << 10) | ((uint32_t)rulings << 11) |
         ((uint32_t)spacing << 14) | ((uint32_t)aperture << 16) | ((uint32_t)aperture_xf << 19) |
         ((uint32_t)aperture_flags << 22);
}

// the kinds of a record, packed as a specialised kernel fixes them
inline uint32_t kinds_of(const optk_surface &s) {
  return kinds_pack(s.transform.kind, s.sag_kind, s.sag_transform.kind, s.material_kind,
                    s.is_mirror != 0, s.rulings_kind, s.spacing_kind, s.aperture_kind,
                    s.aperture_transform.kind, s.aperture_flags);
}

template <uint32_t K> struct Kinds {
  static constexpr bool fixed = (K & KINDS_FIXED) != 0;
  static constexpr int xf = fixed ? (int)(K & 3u) : -1;
  static constexpr int sag = (int)((K >> 2) & 7u);
  static constexpr int sag_xf = fixed ? (int)((K >> 5) & 3u) : -1;
  static constexpr int material = (int)((K >> 7) & 7u);
  static constexpr int mirror = (int)((K >> 10) & 1u);
  static constexpr int rulings = (int)((K >> 11) & 7u);
  static constexpr int spacing = (int)((K >> 14) & 3u);
  static constexpr int aperture = (int)((K >> 16) & 7u);
  static constexpr int aperture_xf = fixed ? (int)((K >> 19) & 3u) : -1;
  static constexpr int aperture_flags = (int)((K >> 22) & 7u);
};

// a record's kind: the compile-time one, or the record's own
template <uint32_t K> OPTK_HD int k_sag(const optk_surface &s) {
  return Kinds<K>::fixed ? Kinds<K>::sag : s.sag_kind;
}
template <uint32_t K> OPTK_HD int k_material(const optk_surface &s) {
  return Kinds<K>::fixed ? Kinds<K>::material : s.material_kind;
}
template <uint32_t K> OPTK_HD bool k_mirror(const optk_surface &s) {
  return Kinds<K>::fixed ? Kinds<K>::mirror != 0 : s.is_mirror != 0;
}
template <uint32_t K> OPTK_HD int k_rulings(const optk_surface &s) {
  return Kinds<K>::fixed ? Kinds<K>::rulings : s.rulings_kind;
}
template <uint32_t K> OPTK_HD int k_spacing(const optk_surface &s) {
  return Kinds<K>::fixed ? Kinds<K>::spacing : s.spacing_kind;
}
template <uint32_t K> OPTK_HD int k_aperture(const optk_surface &s) {
  return Kinds<K>::fixed ? Kinds<K>::aperture : s.aperture_kind;
}
template <uint32_t K> OPTK_HD int k_aperture_flags(const optk_surface &s) {
  return Kinds<K>::fixed ? Kinds<K>::aperture_flags : s.aperture_flags;
}

// Per-surface constants the kernels read instead of recomputing per ray.
// The curvature is the same IEEE quotient the reference forms once per call
// (``c = 1 / self.radius``), so precomputing it changes no result.
OPTK_HD void derive_constants(optk_surface &s) { s.sag_c = 1.0 / s.sag_p[0]; }

struct SystemView {
  const optk_surface *surfaces;
  int num_surfaces;
  int num_paths;
  StackView sv;
};

// ---------------------------------------------------------------------------
// sags (optika/sags/*.py)
// ---------------------------------------------------------------------------

OPTK_HD double sag_z(const optk_surface &s, double x, double y, double z) {
  const optk_affine &T = s.sag_transform;
  if (s.sag_kind == OPTK_SAG_NONE) {  // _flat.py:26-39
    xf_inverse_pos(T, x, y, z);
    z = 0.0;
    xf_apply_pos(T, x, y, z);
    return z;
  }
  xf_inverse_pos(T, x, y, z);
  switch (s.sag_kind) {
  case OPTK_SAG_SPHERE: {  // _spherical.py:100-123
    double c = s.sag_c;
    double r2 = x * x + y * y;
    return c * r2 / (1.0 + sqrt(1.0 - c * c * r2));
  }
  case OPTK_SAG_CONIC:
  case OPTK_SAG_PARABOLA: {  // _conic.py:31-48
    double c = s.sag_c;
    double conic = s.sag_p[1];
    double r2 = x * x + y * y;
    return c * r2 / (1.0 + sqrt(1.0 - (1.0 + conic) * (c * c) * r2));
  }
  case OPTK_SAG_CYLINDER: {  // _cylindrical.py:75-91
    double c = s.sag_c;
    double r2 = x * x;
    return c * r2 / (1.0 + sqrt(1.0 - c * c * r2));
  }
  case OPTK_SAG_TOROID: {  // _toroidal.py:38-57
    double c = s.sag_c;
    double r = s.sag_p[1];
    double x2 = x * x;
    double y2 = y * y;
    double zy = c * y2 / (1.0 + sqrt(1.0 - c * c * y2));
    double rz = r - zy;
    return r - sqrt(rz * rz - x2);
  }
  }
  return 0.0;
}

template <uint32_t K = 0>
OPTK_HD V3<double> sag_normal(const optk_surface &s, double x, double y, double z) {
  if (k_sag<K>(s) == OPTK_SAG_NONE) return {0.0, 0.0, -1.0};  // _flat.py:41-45
  xf_inverse_pos<Kinds<K>::sag_xf>(s.sag_transform, x, y, z);
  switch (k_sag<K>(s)) {
  case OPTK_SAG_SPHERE: {  // _spherical.py:125-146 (not renormalised)
    double c = s.sag_c;
    double nx = c * x;
    double ny = c * y;
    return {nx, ny, -sqrt(1.0 - nx * nx - ny * ny)};
  }
  case OPTK_SAG_CONIC: {  // _conic.py:50-75
    double c = s.sag_c;
    double conic = s.sag_p[1];
    double x2 = x * x;
    double y2 = y * y;
    double c2 = c * c;
    double g = sqrt(1.0 - (1.0 + conic) * c2 * (x2 + y2));
    V3<double> v = {c * x / g, c * y / g, -1.0};
    return div3(v, len3(v));
  }
  case OPTK_SAG_PARABOLA: {  // _parabolic.py:40-63 (numexpr expression)
    double r = s.sag_p[0];
    double xr = x / r, yr = y / r;
    double q = sqrt(xr * xr + yr * yr + 1.0);
    return {x / q / r, y / q / r, -r / q / r};
  }
  case OPTK_SAG_CYLINDER: {  // _cylindrical.py:93-113
    double nx = x / s.sag_p[0];
    return {nx, 0.0, -sqrt(1.0 - nx * nx)};
  }
  case OPTK_SAG_TOROID: {  // _toroidal.py:59-88
    double c = s.sag_c;
    double r = s.sag_p[1];
    double x2 = x * x;
    double y2 = y * y;
    double c2 = c * c;
    double g = sqrt(1.0 - c2 * y2);
    double zy = c * y2 / (1.0 + g);
    double rz = r - zy;
    double f = sqrt(rz * rz - x2);
    double dzdx = x / f;
    double dzydy = c * y / g;
    double dzdy = rz * dzydy / f;
    V3<double> v = {dzdx, dzdy, -1.0};
    return div3(v, len3(v));
  }
  }
  return {0.0, 0.0, -1.0};
}

#define OPTK_NEWTON_MAX_ITER 64
#define OPTK_NEWTON_TOL 1e-13

// Intercept in the sag frame: returns the path length t along u.
template <uint32_t K = 0>
OPTK_HD double sag_intercept_t(const optk_surface &s, double ox, double oy, double oz, double ux,
                               double uy, double uz) {
  switch (k_sag<K>(s)) {
  case OPTK_SAG_NONE:  // _flat.py:47-64
    return -oz / uz;
  case OPTK_SAG_SPHERE: {  // _spherical.py:148-191
    double r = s.sag_p[0];
    double px = ox, py = oy, pz = oz - r;
    double q = ux * px + uy * py + uz * pz;
    return -q - sign_of(r * uz) * sqrt(q * q - (px * px + py * py + pz * pz - r * r));
  }
  case OPTK_SAG_CONIC: {  // _conic.py:120-170
    double c = s.sag_c;
    double kp1 = 1.0 + s.sag_p[1];
    double a = c * (ux * ux + uy * uy + kp1 * (uz * uz));
    double b = 2.0 * (c * (ox * ux + oy * uy + kp1 * oz * uz) - uz);
    double cc = c * (ox * ox + oy * oy + kp1 * (oz * oz)) - 2.0 * oz;
    double disc = b * b - 4.0 * a * cc;
    bool real = disc >= 0.0;
    double sq = sqrt(real ? disc : 0.0);
    bool degenerate = fabs(a) < 1e-12;
    double denom = degenerate ? 1.0 : 2.0 * a;
    double t_lin = 0.0;
    if (degenerate) t_lin = -cc / b;  // only degenerate lanes use it
    double tr[2];
    for (int k = 0; k < 2; ++k) {
      double sgn = k == 0 ? -1.0 : 1.0;
      double t = degenerate ? t_lin : (-b + sgn * sq) / denom;
      double px = ox + ux * t, py = oy + uy * t, pz = oz + uz * t;
      double r2 = px * px + py * py;
      bool on_sheet = (pz * (c * r2 - pz)) >= 0.0;
      tr[k] = (real && on_sheet) ? t : INFINITY;
    }
    return fabs(tr[0]) <= fabs(tr[1]) ? tr[0] : tr[1];
  }
  case OPTK_SAG_PARABOLA: {  // _parabolic.py:142-151 (numexpr expression)
    double f = s.sag_p[2];
    double den = ux * ux + uy * uy;
    double oyux = oy * ux, oxuy = ox * uy;
    double tq = (-ox * ux - oy * uy + 2.0 * f * uz -
                 sign_of(f * uz) *
                     sqrt(-(oyux * oyux) - oxuy * oxuy + 2.0 * oy * uy * (ox * ux - 2.0 * f * uz) +
                          4.0 * f * (oz * (ux * ux + uy * uy) - ox * ux * uz + f * (uz * uz)))) /
                (ux * ux + uy * uy);
    double tl = (ox * ox + oy * oy - 4.0 * f * oz) / (4.0 * f * uz);
    return den > 1e-10 ? tq : tl;
  }
  case OPTK_SAG_CYLINDER: {  // _cylindrical.py:115-160
    double r = s.sag_p[0];
    double bx = -ox, by = -oy, bz = r - oz;
    double ncx = -uz, ncy = 0.0, ncz = ux;  // u x (0, 1, 0)
    double nca2 = ncx * ncx + ncy * ncy + ncz * ncz;
    double bcx = -bz, bcy = 0.0, bcz = bx;  // b x (0, 1, 0)
    double negative_b = ncx * bcx + ncy * bcy + ncz * bcz;
    double b_squared = nca2 * (r * r);
    double bn = bx * ncx + by * ncy + bz * ncz;
    double disc = b_squared - bn * bn;
    double sgn = sign_of(r * uz);
    return disc > 0.0 ? (negative_b - sgn * sqrt(disc)) / nca2 : -oz / uz;
  }
  }
  return -oz / uz;
}

// Toroid intercept: AbstractSag.intercept (optika/sags/_abc.py:76-107) solves
// line(t).z - sag(line(t)) = 0 along the ray in the *surface* frame, the sag
// transformation acting inside sag() only.  Newton iteration from t = 0.
template <uint32_t K = 0>
OPTK_HD double toroid_intercept_t(const optk_surface &s, double ox, double oy, double oz, double ux,
                                  double uy, double uz) {
  const optk_affine &T = s.sag_transform;
  double c = s.sag_c;
  double r = s.sag_p[1];
  double vx = ux, vy = uy, vz = uz;  // direction in the sag frame (chain rule)
  xf_inverse_dir<Kinds<K>::sag_xf>(T, vx, vy, vz);
  double t = 0.0;
  for (int it = 0; it < OPTK_NEWTON_MAX_ITER; ++it) {
    double x = ox + ux * t;
    double y = oy + uy * t;
    double zl = oz + uz * t;
    xf_inverse_pos<Kinds<K>::sag_xf>(T, x, y, zl);
    double x2 = x * x, y2 = y * y;
    double g = sqrt(1.0 - c * c * y2);
    double zy = c * y2 / (1.0 + g);
    double rz = r - zy;
    double f = sqrt(rz * rz - x2);
    double z = r - f;
    double dzdx = x / f;
    double dzdy = rz * (c * y / g) / f;
    double fval = (oz + uz * t) - z;
    double fder = uz - (dzdx * vx + dzdy * vy);
    double step = fval / fder;
    t = t - step;
    if (fabs(step) <= OPTK_NEWTON_TOL) break;
  }
  return t;
}

// AbstractSag.propagate_rays (optika/sags/_abc.py:109-122) in the surface
// frame: move to the intercept and attenuate.
template <int F = FEAT_ALL, uint32_t K = 0>
OPTK_HD void sag_propagate(const optk_surface &s, Ray<double> &r) {
  double nx, ny, nz;
  if ((F & FEAT_TOROID) && k_sag<K>(s) == OPTK_SAG_TOROID) {
    double t = toroid_intercept_t<K>(s, r.px, r.py, r.pz, r.dx, r.dy, r.dz);
    nx = r.px + r.dx * t;
    ny = r.py + r.dy * t;
    nz = r.pz + r.dz * t;
  } else {
    const optk_affine &T = s.sag_transform;
    double ox = r.px, oy = r.py, oz = r.pz;
    double ux = r.dx, uy = r.dy, uz = r.dz;
    xf_inverse_ray<Kinds<K>::sag_xf>(T, ox, oy, oz, ux, uy, uz);
    double t = sag_intercept_t<K>(s, ox, oy, oz, ux, uy, uz);
    nx = ox + ux * t;
    ny = oy + uy * t;
    nz = oz + uz * t;
    xf_apply_ray<Kinds<K>::sag_xf>(T, nx, ny, nz, ux, uy, uz);
    r.dx = ux;
    r.dy = uy;
    r.dz = uz;
  }
  double ddx = nx - r.px, ddy = ny - r.py, ddz = nz - r.pz;
  // exp(-alpha |d|) is exactly 1 for alpha = 0 and a finite displacement:
  // skip the square root and the exponential for non-attenuating media
  double f = 1.0;
  if (!(r.attenuation == 0.0 && fabs(ddx) < 1e150 && fabs(ddy) < 1e150 && fabs(ddz) < 1e150)) {
    double len = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
    f = exp(-r.attenuation * len);
  }
  r.intensity = f * r.intensity;
  r.px = nx;
  r.py = ny;
  r.pz = nz;
}

// ---------------------------------------------------------------------------
// rulings (optika/rulings/_rulings.py, _spacing.py)
// ---------------------------------------------------------------------------

// spacing_(position, normal): the local ruling vector d g
template <typename T, uint32_t K = 0>
OPTK_HD V3<T> spacing_vector(const optk_surface &s, V3<T> p, V3<T> n) {
  switch (k_spacing<K>(s)) {
  case OPTK_SPACING_CONSTANT:  // _spacing.py:69-74
    return {T(s.spacing_constant) * T(s.spacing_normal[0]),
            T(s.spacing_constant) * T(s.spacing_normal[1]),
            T(s.spacing_constant) * T(s.spacing_normal[2])};
  case OPTK_SPACING_POLY1D: {  // _spacing.py:109-128
    xf_apply_pos(s.spacing_transform, p.x, p.y, p.z);
    V3<T> g = {T(s.spacing_normal[0]), T(s.spacing_normal[1]), T(s.spacing_normal[2])};
    T x = dot3(p, g);
    T d = T(0);
    for (int k = 0; k < s.num_terms; ++k) {
      int pw = s.poly_power[k];
      T xp;
      if (pw == 0)
        xp = T(1);
      else if (pw == 1)
        xp = x;
      else if (pw == 2)
        xp = x * x;
      else
        xp = pow(x, (T)pw);
      d = d + T(s.poly_coeff[k]) * xp;
    }
    return {d * g.x, d * g.y, d * g.z};
  }
  case OPTK_SPACING_HOLOGRAPHIC: {  // _spacing.py:295-328
    T d1 = T(2) * (T)s.holo_diverging1 - T(1);
    T d2 = T(2) * (T)s.holo_diverging2 - T(1);
    V3<T> r1 = {p.x - T(s.holo_x1[0]), p.y - T(s.holo_x1[1]), p.z - T(s.holo_x1[2])};
    V3<T> r2 = {p.x - T(s.holo_x2[0]), p.y - T(s.holo_x2[1]), p.z - T(s.holo_x2[2])};
    T l1 = len3(r1), l2 = len3(r2);
    r1 = {d1 * (r1.x / l1), d1 * (r1.y / l1), d1 * (r1.z / l1)};
    r2 = {d2 * (r2.x / l2), d2 * (r2.y / l2), d2 * (r2.z / l2)};
    V3<T> dr = {r1.x - r2.x, r1.y - r2.y, r1.z - r2.z};
    V3<T> aq = cross3(n, dr);
    T a = len3(aq);
    V3<T> q = div3(aq, a);
    T spacing = T(s.holo_wavelength) / a;
    V3<T> qn = cross3(q, n);
    return {spacing * qn.x, spacing * qn.y, spacing * qn.z};
  }
  }
  return {T(0), T(0), T(0)};
}

OPTK_HD int path_order(const optk_surface &s, int path) {
  return s.num_orders > 1 ? s.orders[path] : s.orders[0];
}

// incident_effective (_rulings.py:24-128, 170-204)
OPTK_HD V3<double> incident_effective(const optk_surface &s, int order, double wavelength,
                                      V3<double> a, double n, V3<double> normal,
                                      V3<double> kappa) {
  double spacing = len3(kappa);
  V3<double> g = div3(kappa, spacing);
  double w = wavelength * s.wavelength_scale;
  double sgn = sign_of(dot3(a, normal));
  double m = (double)order;
  double den = n * spacing;
  return {a.x + sgn * m * w * g.x / den, a.y + sgn * m * w * g.y / den,
          a.z + sgn * m * w * g.z / den};
}

// Bessel function of the first kind of integer order, per real type
OPTK_HD double bessel_jn(int m, double x) { return jn(m, x); }
OPTK_HD float bessel_jn(int m, float x) { return jnf(m, x); }

// linear interpolation in a (wavelength row, value row) table, per real type
OPTK_HD double table_interp(const double *w, int count, double x) {
  return interp1(w, w + count, count, x);
}

OPTK_HD float table_interp(const float *w, int count, float x) {
  return interp1f(w, w + count, count, x);
}

// groove efficiencies (_rulings.py:246-1073)
template <typename T, uint32_t K = 0>
OPTK_HD T rulings_efficiency(const optk_surface &s, int order, T wavelength, V3<T> a_eff,
                             V3<T> position, V3<T> normal, const T *tables) {
  switch (k_rulings<K>(s)) {
  case OPTK_RULINGS_IDEAL:
    return T(1);
  case OPTK_RULINGS_MEASURED:
    return table_interp(tables + s.rulings_table, s.rulings_table_len, wavelength);
  default:
    break;
  }
  V3<T> kappa = spacing_vector<T, K>(s, position, normal);
  V3<T> nr = div3(kappa, len3(kappa));
  V3<T> par = cross3(normal, nr);
  par = div3(par, len3(par));
  T proj = dot3(a_eff, par);
  V3<T> d = {a_eff.x - proj, a_eff.y - proj, a_eff.z - proj};
  T cos_theta = -dot3(d, normal);
  T i = (T)order;
  const T pi = T(M_PI);
  const T depth = T(s.rulings_p[0]);
  switch (k_rulings<K>(s)) {
  case OPTK_RULINGS_SINUSOIDAL: {  // J_i(2 gamma), unsquared
    T gamma = pi * depth / (wavelength * cos_theta);
    // J_{-n}(x) = (-1)^n J_n(x), applied explicitly: the device Bessel
    // functions are only relied on for n >= 0
    int m = order < 0 ? -order : order;
    T v = bessel_jn(m, T(2) * gamma);
    return (order < 0 && (m & 1)) ? -v : v;
  }
  case OPTK_RULINGS_SQUARE: {
    T dd = depth / (pi / T(4));
    T gamma = pi * dd / (wavelength * cos_theta);
    if (order == 0) {
      T cg = cos(pi * gamma / T(2));
      return cg * cg;
    }
    if ((order & 1) == 0) return T(0);
    T v = T(2) * sin(pi * gamma / T(2)) / (i * pi);
    return v * v;
  }
  case OPTK_RULINGS_SAWTOOTH: {
    T dd = depth / (pi / T(2));
    T gamma = pi * dd / (wavelength * cos_theta);
    T v = sin(pi * gamma) / (pi * (gamma + i));
    return v * v;
  }
  case OPTK_RULINGS_TRIANGULAR: {
    T dd = depth / ((pi * pi) / T(8));
    T gamma = pi * dd / (wavelength * cos_theta);
    T pg = pi * gamma / T(2);
    T a = gamma / (pg * pg + i * i);
    T arg = (pi * pi) * gamma / T(4);
    T v = (order & 1) == 0 ? a * sin(arg) : a * cos(arg);
    return v * v;
  }
  case OPTK_RULINGS_RECTANGULAR: {
    T a = T(2) * pi * T(s.rulings_p[1]);
    T root = sqrt(T(2) * (T(1) - cos(a)));
    T amplitude = pi / (T(2) * root);
    T dd = depth / amplitude;
    T gamma = pi * dd / (wavelength * cos_theta);
    T sb = sin(pi * gamma / root);
    T b = sb * sb;
    if (order == 0) {
      T ap = a / pi;
      return T(1) - ((T(2) * a / pi) - ap * ap) * b;
    }
    T ip = i * pi;
    return (T(2) / (ip * ip)) * (T(1) - cos(i * a)) * b;
  }
  }
  return T(1);
}

// ---------------------------------------------------------------------------
// materials (optika/materials/_materials.py, _multilayers.py)
// ---------------------------------------------------------------------------

template <uint32_t K = 0>
OPTK_HD double material_index(const optk_surface &s, double wavelength, double n1) {
  switch (k_material<K>(s)) {
  case OPTK_MATERIAL_VACUUM:
    return 1.0;
  case OPTK_MATERIAL_GLASS: {  // _materials.py:428-438
    double w2 = wavelength * wavelength;
    const double *p = s.material_p;
    double n2 = 1.0 + (p[0] * w2 / (w2 - p[3]) + p[1] * w2 / (w2 - p[4]) + p[2] * w2 / (w2 - p[5]));
    return sqrt(n2);
  }
  default:
    return n1;
  }
}

template <uint32_t K = 0>
OPTK_HD double material_attenuation(const optk_surface &s, double attenuation) {
  const int kind = k_material<K>(s);
  return (kind == OPTK_MATERIAL_VACUUM || kind == OPTK_MATERIAL_GLASS) ? 0.0 : attenuation;
}

// BackIlluminatedSiliconSensorMaterial.efficiency (optika/sensors/materials/
// _materials.py:2231-2243): the s/p average of absorbance(...,
// method="Beer-Lambert") (_materials.py:200-376) -- the transmittance
// through the oxide into the silicon times 1 - exp(-alpha d), with
// alpha = 4 pi Im(n_sub cos_sub) / wavelength the absorption per unit
// perpendicular depth of the refracted wave (absorption_effective,
// _materials.py:55-93), in the reference's operation order.
OPTK_HD double silicon_absorbance_average(const StackView &sv, int stack, double thickness,
                                          double wavelength, C<double> direction,
                                          C<double> n) {
  MultilayerRT<double> rt = multilayer_efficiency<double>(sv, stack, wavelength, direction, n, true);
  const optk_stack &stk = sv.stacks[stack];
  C<double> n_sub = {1.0, 0.0};
  if (stk.substrate_layer >= 0) n_sub = layer_index<double>(sv.layers[stk.substrate_layer], sv, wavelength);
  C<double> d_sub = snell_scalar(direction, n, n_sub);
  double absorption = 4.0 * M_PI * cmul(n_sub, d_sub).im / wavelength;
  double decay = exp(-absorption * thickness);
  double s_abs = rt.T[0] - rt.T[0] * decay;
  double p_abs = rt.T[1] - rt.T[1] * decay;
  return (s_abs + p_abs) / 2.0;
}

template <int F = FEAT_ALL, uint32_t K = 0>
OPTK_HD double material_efficiency(const optk_surface &s, const StackView &sv, double wavelength,
                                   V3<double> a, double index, double attenuation,
                                   V3<double> normal) {
  switch (k_material<K>(s)) {
  case OPTK_MATERIAL_MEASURED_MIRROR: {  // _materials.py:279-305
    const double *w = sv.tables + s.material_table;
    return interp1(w, w + s.material_table_len, s.material_table_len, wavelength);
  }
  case OPTK_MATERIAL_MULTILAYER_MIRROR:
  case OPTK_MATERIAL_MULTILAYER_FILM: {  // _multilayers.py:839-935
    if (!(F & FEAT_MULTILAYER)) return 1.0;
    double k = attenuation * wavelength / (4.0 * M_PI) * 1e-6;
    C<double> n = {index, k};
    double cosv = -dot3(a, normal);
    bool film = k_material<K>(s) == OPTK_MATERIAL_MULTILAYER_FILM;
    MultilayerRT<double> rt =
        multilayer_efficiency<double>(sv, s.stack, wavelength, C<double>{cosv, 0.0}, n, film);
    return film ? (rt.T[0] + rt.T[1]) / 2.0 : (rt.R[0] + rt.R[1]) / 2.0;
  }
  case OPTK_MATERIAL_SILICON_SENSOR: {  // _materials.py:2231-2243
    if (!(F & FEAT_MULTILAYER)) return 1.0;
    double k = attenuation * wavelength / (4.0 * M_PI) * 1e-6;
    C<double> n = {index, k};
    double cosv = -dot3(a, normal);
    return silicon_absorbance_average(sv, s.stack, s.material_p[0], wavelength,
                                      C<double>{cosv, 0.0}, n);
  }
  default:
    return 1.0;
  }
}

// vector Snell's law (optika/materials/_snells_law.py:341-366)
// n1 / n2, with the (exact) value 1 when the indices agree: mirrors and
// vacuum-to-vacuum surfaces then skip two float64 divisions
OPTK_HD double index_ratio(double n1, double n2) {
  return (n1 == n2 && n1 != 0.0 && fabs(n1) < INFINITY) ? 1.0 : n1 / n2;
}

OPTK_HD V3<double> snells_law(V3<double> a, double n1, double n2, V3<double> u, bool is_mirror) {
  double a2 = a.x * a.x + a.y * a.y + a.z * a.z;
  double r = index_ratio(n1, n2);
  double r2 = r * r;
  double au = a.x * u.x + a.y * u.y + a.z * u.z;
  double au2 = au * au;
  double sgn = -copysign(1.0, au);
  double inv_r2 = r == 1.0 ? 1.0 : 1.0 / r2;
  double d = -au + sgn * (2.0 * (double)is_mirror - 1.0) * sqrt(inv_r2 + au2 - a2);
  return {r * (a.x + d * u.x), r * (a.y + d * u.y), r * (a.z + d * u.z)};
}

// ---------------------------------------------------------------------------
// apertures (optika/apertures/_apertures.py)
// ---------------------------------------------------------------------------

template <uint32_t K = 0>
OPTK_HD bool aperture_contains(const optk_surface &s, const double *tables, double x, double y,
                               double z) {
  xf_inverse_pos<Kinds<K>::aperture_xf>(s.aperture_transform, x, y, z);
  bool mask = true;
  switch (k_aperture<K>(s)) {
  case OPTK_APERTURE_CIRCLE:  // _apertures.py:292-314
    mask = sqrt(x * x + y * y) <= s.aperture_p[0];
    break;
  case OPTK_APERTURE_SECTOR: {  // _apertures.py:438-480
    bool mr = sqrt(x * x + y * y) <= s.aperture_p[0];
    double ang = atan2(y, x);
    const double two_pi = 2.0 * M_PI;
    double apos = fmod(ang, two_pi);
    if (apos != 0.0 && apos < 0.0) apos += two_pi;  // numpy remainder, divisor > 0
    double aneg = fmod(ang, -two_pi);
    if (aneg != 0.0 && aneg > 0.0) aneg -= two_pi;  // numpy remainder, divisor < 0
    double a0 = s.aperture_p[1], a1 = s.aperture_p[2];
    mask = mr && (((a0 < apos) && (apos < a1)) || ((a0 < aneg) && (aneg < a1)));
    break;
  }
  case OPTK_APERTURE_ELLIPSE: {  // _apertures.py:640-663
    double u = x / s.aperture_p[0], v = y / s.aperture_p[1];
    mask = u * u + v * v <= 1.0;
    break;
  }
  case OPTK_APERTURE_RECTANGLE: {  // _apertures.py:941-968
    double hx = s.aperture_p[0], hy = s.aperture_p[1];
    mask = ((-hx <= x) && (x <= hx)) && ((-hy <= y) && (y <= hy));
    break;
  }
  case OPTK_APERTURE_POLYGON: {  // crossing-number rule
    const double *vx = tables + s.polygon_offset;
    const double *vy = vx + s.polygon_count;
    int nv = s.polygon_count;
    bool inside = false;
    int j = nv - 1;
    for (int i = 0; i < nv; ++i) {
      double xi = vx[i], yi = vy[i], xj = vx[j], yj = vy[j];
      bool cond = (yi > y) != (yj > y);
      double xc = (xj - xi) * (y - yi) / (yj - yi) + xi;
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

// ---------------------------------------------------------------------------
// the surface step
// ---------------------------------------------------------------------------

template <int F = FEAT_ALL, uint32_t K = 0>
OPTK_HD void propagate_surface(const optk_surface &s, const SystemView &sys, int path,
                               Ray<double> &r) {
  xf_inverse_ray<Kinds<K>::xf>(s.transform, r.px, r.py, r.pz, r.dx, r.dy, r.dz);

  sag_propagate<F, K>(s, r);
  V3<double> p1 = {r.px, r.py, r.pz};
  V3<double> normal = sag_normal<K>(s, p1.x, p1.y, p1.z);

  V3<double> a = {r.dx, r.dy, r.dz};
  int order = 0;
  const bool ruled = (F & FEAT_RULINGS) && k_rulings<K>(s) != OPTK_RULINGS_NONE;
  if (ruled) {
    order = path_order(s, path);
    V3<double> kappa = spacing_vector<double, K>(s, p1, normal);
    a = incident_effective(s, order, r.wavelength, a, r.index, normal, kappa);
  }

  double n1 = r.index;
  double n2 = material_index<K>(s, r.wavelength, n1);
  double ratio = index_ratio(n1, n2);
  double wavelength_2 = ratio == 1.0 ? r.wavelength : r.wavelength / ratio;
  V3<double> b = snells_law(a, n1, n2, normal, k_mirror<K>(s));
  double eff =
      material_efficiency<F, K>(s, sys.sv, r.wavelength, a, r.index, r.attenuation, normal);
  if (ruled)
    eff = eff * rulings_efficiency<double, K>(s, order, r.wavelength, a, p1, normal, sys.sv.tables);
  r.intensity = r.intensity * eff;
  r.attenuation = material_attenuation<K>(s, r.attenuation);
  r.index = n2;
  r.wavelength = wavelength_2;
  r.dx = b.x;
  r.dy = b.y;
  r.dz = b.z;

  if (k_aperture<K>(s) != OPTK_APERTURE_NONE) {
    bool m = (k_aperture_flags<K>(s) & OPTK_AP_DIRECTION)
                 ? aperture_contains<K>(s, sys.sv.tables, r.dx, r.dy, r.dz)
                 : aperture_contains<K>(s, sys.sv.tables, r.px, r.py, r.pz);
    r.unvignetted = r.unvignetted && m;
  }

  xf_apply_ray<Kinds<K>::xf>(s.transform, r.px, r.py, r.pz, r.dx, r.dy, r.dz);
}

// ---------------------------------------------------------------------------
// system-level driver pieces shared by the kernels and the host test build
// ---------------------------------------------------------------------------

struct TraceSpec {
  SystemView sys;
  optk_affine pre[2];
  int32_t pre_is_inverse[2];
  optk_affine post;
  int32_t apply_post;
};

// the object's transformation (forward) then the system's (inverse)
OPTK_HD void apply_pre(const TraceSpec &P, Ray<double> &r) {
  for (int k = 0; k < 2; ++k) {
    const optk_affine &a = P.pre[k];
    if (a.kind == OPTK_XF_IDENTITY) continue;
    if (P.pre_is_inverse[k]) {
      xf_inverse_ray(a, r.px, r.py, r.pz, r.dx, r.dy, r.dz);
    } else {
      xf_apply_ray(a, r.px, r.py, r.pz, r.dx, r.dy, r.dz);
    }
  }
}

OPTK_HD void apply_post(const TraceSpec &P, Ray<double> &r) {
  if (!P.apply_post) return;
  xf_inverse_ray(P.post, r.px, r.py, r.pz, r.dx, r.dy, r.dz);
}

// sin and cos of |x| <= 0.01 rad by their Maclaurin series through x^7 /
// x^8 in nested form (multiplies and adds, no fused operations): the next
// terms, x^9 / 9! and x^10 / 10!, are below 1e-21 relative, so the result
// is the exact value to within the rounding of the few operations (within
// 1 ulp of numpy's sin and cos over the range, pinned by the tests)
OPTK_HD void small_angle_sincos(double x, double &s, double &c) {
  const double x2 = x * x;
  s = x * (1.0 - x2 * (1.0 / 6.0) * (1.0 - x2 * (1.0 / 20.0) * (1.0 - x2 * (1.0 / 42.0))));
  c = 1.0 - x2 * 0.5 * (1.0 - x2 * (1.0 / 12.0) * (1.0 - x2 * (1.0 / 30.0) * (1.0 - x2 * (1.0 / 56.0))));
}

// optika.direction (optika/_util.py:41-73): R_y(phi_x) R_x(phi_y) z.  On the
// device the field angles of imaging systems (|angle| <= 0.01 rad, 0.57
// deg) take the short series above instead of the library's range-reduced
// sincos, whose reduction and quadrant selection cost several times more
// instructions; larger angles take the library.  Both agree with numpy's
// sin and cos to a few ulp.
OPTK_HD void direction_from_angles(double ax_deg, double ay_deg, double &dx, double &dy,
                                   double &dz) {
  const double deg = 0.017453292519943295;  // numpy.deg2rad factor pi / 180
  double ax = ax_deg * deg, ay = ay_deg * deg;
  double sx, cx, sy, cy;
#if defined(__HIP_DEVICE_COMPILE__)
  if (fabs(ax) <= 0.01 && fabs(ay) <= 0.01) {
    small_angle_sincos(ax, sx, cx);
    small_angle_sincos(ay, sy, cy);
  } else {
    sincos(ax, &sx, &cx);
    sincos(ay, &sy, &cy);
  }
#else
  sx = sin(ax);
  cx = cos(ax);
  sy = sin(ay);
  cy = cos(ay);
#endif
  dx = -cy * sx;
  dy = -sy;
  dz = cy * cx;
}

// _calc_rayfunction_input (optika/systems/_sequential.py:791-828) for
// physical field/pupil coordinates
OPTK_HD Ray<double> grid_ray(double fx, double fy, double ux, double uy, double wavelength,
                             double intensity, int object_at_infinity) {
  Ray<double> r;
  if (object_at_infinity) {
    r.px = ux;
    r.py = uy;
    direction_from_angles(fx, fy, r.dx, r.dy, r.dz);
  } else {
    r.px = fx;
    r.py = fy;
    direction_from_angles(ux, uy, r.dx, r.dy, r.dz);
  }
  r.pz = 0.0;
  r.wavelength = wavelength;
  r.intensity = intensity;
  r.attenuation = 0.0;
  r.index = 1.0;
  r.unvignetted = true;
  return r;
}

}  // namespace optk
