// optk_physics.h -- per-ray physics of the sequential trace, shared by the
// gfx950 kernels (hipcc) and a host build used only by the CPU test harness.
//
// One thread owns one ray.  Everything below is straight-line per-ray code
// over a uniform surface record: branches on record fields are uniform
// across a wavefront (the compiler keeps the record in scalar registers),
// branches on ray data become selects.
//
// Float64 parity with the reference's numpy arithmetic: expressions keep the
// reference's operation order, the library is compiled with
// -ffp-contract=off, and complex arithmetic follows numpy's loops
// (componentwise products, Smith's division) so that the float64 kernel
// agrees with a numpy float64 evaluation to rounding (correctly rounded
// +,-,*,/,sqrt; transcendental functions to within an ulp).
#pragma once

#include <math.h>
#include <stdint.h>
#include <string.h>

#include "../../include/optk.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define OPTK_HD __host__ __device__ __forceinline__
#else
#define OPTK_HD inline
#endif

namespace optk {

// ---------------------------------------------------------------------------
// real helpers
// ---------------------------------------------------------------------------

template <typename T> OPTK_HD T sign_of(T x) {  // numpy.sign
  return x > T(0) ? T(1) : (x < T(0) ? T(-1) : (x == T(0) ? T(0) : x));
}

template <typename T> struct V3 {
  T x, y, z;
};

template <typename T> OPTK_HD T dot3(const V3<T> &a, const V3<T> &b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <typename T> OPTK_HD T len3(const V3<T> &a) {
  return sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
}
template <typename T> OPTK_HD V3<T> cross3(const V3<T> &a, const V3<T> &b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename T> OPTK_HD V3<T> div3(const V3<T> &a, T s) {
  return {a.x / s, a.y / s, a.z / s};
}

// ---------------------------------------------------------------------------
// complex arithmetic with numpy's semantics
// ---------------------------------------------------------------------------

template <typename T> struct C {
  T re, im;
};

template <typename T> OPTK_HD C<T> cadd(C<T> a, C<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T> OPTK_HD C<T> csub(C<T> a, C<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T> OPTK_HD C<T> cmul(C<T> a, C<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// real * complex (numpy promotes the real operand; the zero imaginary part
// contributes exact zeros for finite values)
template <typename T> OPTK_HD C<T> cscale(T a, C<T> b) { return {a * b.re, a * b.im}; }
template <typename T> OPTK_HD C<T> cconj(C<T> a) { return {a.re, -a.im}; }

// Smith's division, as numpy's complex divide loop.
template <typename T> OPTK_HD C<T> cdiv(C<T> a, C<T> b) {
  T br_abs = fabs(b.re), bi_abs = fabs(b.im);
  if (br_abs >= bi_abs) {
    if (br_abs == T(0) && bi_abs == T(0)) {
      return {a.re / br_abs, a.im / br_abs};
    }
    T rat = b.im / b.re;
    T scl = T(1) / (b.re + b.im * rat);
    return {(a.re + a.im * rat) * scl, (a.im - a.re * rat) * scl};
  }
  T rat = b.re / b.im;
  T scl = T(1) / (b.im + b.re * rat);
  return {(a.re * rat + a.im) * scl, (a.im * rat - a.re) * scl};
}

template <typename T> OPTK_HD T cabs(C<T> a) { return hypot(a.re, a.im); }

// Principal square root (C99 csqrt algorithm: hypot-based, cancellation-free
// branch selection; the extreme-magnitude rescaling branches are omitted,
// the trace never reaches them).
template <typename T> OPTK_HD C<T> csqrt(C<T> z) {
  T x = z.re, y = z.im;
  if (y == T(0)) {
    if (x < T(0)) return {T(0), copysign(sqrt(-x), y)};
    return {fabs(sqrt(x)), copysign(T(0), y)};
  }
  if (x == T(0)) {
    T r = sqrt(T(0.5) * fabs(y));
    return {r, copysign(r, y)};
  }
  T d = hypot(x, y);
  T r, s;
  if (x > T(0)) {
    r = sqrt(T(0.5) * (d + x));
    s = T(0.5) * (y / r);
  } else {
    s = sqrt(T(0.5) * (d - x));
    r = fabs(T(0.5) * (y / s));
  }
  return {r, copysign(s, y)};
}

// exp(x + i y) = e^x (cos y + i sin y)
template <typename T> OPTK_HD C<T> cexp(C<T> z) {
  T e = exp(z.re);
  T s, c;
#if defined(__HIP_DEVICE_COMPILE__)
  sincos(z.im, &s, &c);
#else
  s = sin(z.im);
  c = cos(z.im);
#endif
  return {e * c, e * s};
}

template <typename T, typename U> OPTK_HD C<T> to_c_impl(C<U> z) { return {T(z.re), T(z.im)}; }
template <typename T> OPTK_HD C<T> to_c(C<double> z) { return to_c_impl<T, double>(z); }

template <typename T> struct M2 {
  C<T> a, b, c, d;  // [[a, b], [c, d]]
};

template <typename T> OPTK_HD M2<T> mmul(const M2<T> &x, const M2<T> &y) {
  return {cadd(cmul(x.a, y.a), cmul(x.b, y.c)), cadd(cmul(x.a, y.b), cmul(x.b, y.d)),
          cadd(cmul(x.c, y.a), cmul(x.d, y.c)), cadd(cmul(x.c, y.b), cmul(x.d, y.d))};
}
template <typename T> OPTK_HD M2<T> midentity() {
  return {{T(1), T(0)}, {T(0), T(0)}, {T(0), T(0)}, {T(1), T(0)}};
}
template <typename T> OPTK_HD M2<T> mselect(bool cond, const M2<T> &x, const M2<T> &y) {
  return cond ? x : y;
}

// ---------------------------------------------------------------------------
// transformations (XK: the record's kind fixed at compile time by a
// specialised kernel, or -1 to read it from the record)
// ---------------------------------------------------------------------------

template <int XK = -1, typename T> OPTK_HD void xf_apply_pos(const optk_affine &a, T &x, T &y, T &z) {
  const int kind = XK >= 0 ? XK : a.kind;
  if (kind == OPTK_XF_IDENTITY) return;
  if (kind == OPTK_XF_AFFINE) {
    T X = T(a.m[0]) * x + T(a.m[1]) * y + T(a.m[2]) * z;
    T Y = T(a.m[3]) * x + T(a.m[4]) * y + T(a.m[5]) * z;
    T Z = T(a.m[6]) * x + T(a.m[7]) * y + T(a.m[8]) * z;
    x = X;
    y = Y;
    z = Z;
  }
  x = x + T(a.t[0]);
  y = y + T(a.t[1]);
  z = z + T(a.t[2]);
}
template <int XK = -1, typename T> OPTK_HD void xf_apply_dir(const optk_affine &a, T &x, T &y, T &z) {
  const int kind = XK >= 0 ? XK : a.kind;
  if (kind != OPTK_XF_AFFINE) return;
  T X = T(a.m[0]) * x + T(a.m[1]) * y + T(a.m[2]) * z;
  T Y = T(a.m[3]) * x + T(a.m[4]) * y + T(a.m[5]) * z;
  T Z = T(a.m[6]) * x + T(a.m[7]) * y + T(a.m[8]) * z;
  x = X;
  y = Y;
  z = Z;
}
template <int XK = -1, typename T> OPTK_HD void xf_inverse_pos(const optk_affine &a, T &x, T &y, T &z) {
  const int kind = XK >= 0 ? XK : a.kind;
  if (kind == OPTK_XF_IDENTITY) return;
  x = x - T(a.t[0]);
  y = y - T(a.t[1]);
  z = z - T(a.t[2]);
  if (kind == OPTK_XF_AFFINE) {
    T X = T(a.m[0]) * x + T(a.m[3]) * y + T(a.m[6]) * z;
    T Y = T(a.m[1]) * x + T(a.m[4]) * y + T(a.m[7]) * z;
    T Z = T(a.m[2]) * x + T(a.m[5]) * y + T(a.m[8]) * z;
    x = X;
    y = Y;
    z = Z;
  }
}
template <int XK = -1, typename T> OPTK_HD void xf_inverse_dir(const optk_affine &a, T &x, T &y, T &z) {
  const int kind = XK >= 0 ? XK : a.kind;
  if (kind != OPTK_XF_AFFINE) return;
  T X = T(a.m[0]) * x + T(a.m[3]) * y + T(a.m[6]) * z;
  T Y = T(a.m[1]) * x + T(a.m[4]) * y + T(a.m[7]) * z;
  T Z = T(a.m[2]) * x + T(a.m[5]) * y + T(a.m[8]) * z;
  x = X;
  y = Y;
  z = Z;
}

// a ray's position and direction under one record with one branch on its
// kind; the arithmetic of each component is that of xf_apply_pos /
// xf_apply_dir (or xf_inverse_pos / xf_inverse_dir), operation for operation
template <int XK = -1, typename T>
OPTK_HD void xf_apply_ray(const optk_affine &a, T &px, T &py, T &pz, T &dx, T &dy, T &dz) {
  const int kind = XK >= 0 ? XK : a.kind;
  if (kind == OPTK_XF_IDENTITY) return;
  if (kind == OPTK_XF_AFFINE) {
    T X = T(a.m[0]) * px + T(a.m[1]) * py + T(a.m[2]) * pz;
    T Y = T(a.m[3]) * px + T(a.m[4]) * py + T(a.m[5]) * pz;
    T Z = T(a.m[6]) * px + T(a.m[7]) * py + T(a.m[8]) * pz;
    T DX = T(a.m[0]) * dx + T(a.m[1]) * dy + T(a.m[2]) * dz;
    T DY = T(a.m[3]) * dx + T(a.m[4]) * dy + T(a.m[5]) * dz;
    T DZ = T(a.m[6]) * dx + T(a.m[7]) * dy + T(a.m[8]) * dz;
    px = X;
    py = Y;
    pz = Z;
    dx = DX;
    dy = DY;
    dz = DZ;
  }
  px = px + T(a.t[0]);
  py = py + T(a.t[1]);
  pz = pz + T(a.t[2]);
}
template <int XK = -1, typename T>
OPTK_HD void xf_inverse_ray(const optk_affine &a, T &px, T &py, T &pz, T &dx, T &dy, T &dz) {
  const int kind = XK >= 0 ? XK : a.kind;
  if (kind == OPTK_XF_IDENTITY) return;
  px = px - T(a.t[0]);
  py = py - T(a.t[1]);
  pz = pz - T(a.t[2]);
  if (kind == OPTK_XF_AFFINE) {
    T X = T(a.m[0]) * px + T(a.m[3]) * py + T(a.m[6]) * pz;
    T Y = T(a.m[1]) * px + T(a.m[4]) * py + T(a.m[7]) * pz;
    T Z = T(a.m[2]) * px + T(a.m[5]) * py + T(a.m[8]) * pz;
    T DX = T(a.m[0]) * dx + T(a.m[3]) * dy + T(a.m[6]) * dz;
    T DY = T(a.m[1]) * dx + T(a.m[4]) * dy + T(a.m[7]) * dz;
    T DZ = T(a.m[2]) * dx + T(a.m[5]) * dy + T(a.m[8]) * dz;
    px = X;
    py = Y;
    pz = Z;
    dx = DX;
    dy = DY;
    dz = DZ;
  }
}

// ---------------------------------------------------------------------------
// ray state
// ---------------------------------------------------------------------------

template <typename T> struct Ray {
  T px, py, pz;
  T dx, dy, dz;
  T wavelength, intensity, attenuation, index;
  bool unvignetted;
};

// ---------------------------------------------------------------------------
// 1-D linear interpolation with numpy.interp's rules (clamped ends, exact
// knots), over a table of `count` ascending abscissae.
// ---------------------------------------------------------------------------

OPTK_HD int interp_bracket(const double *xp, int count, double x) {
  // index j with xp[j] <= x < xp[j+1]; assumes xp[0] <= x < xp[count-1]
  int lo = 0, hi = count - 1;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (x < xp[mid])
      hi = mid;
    else
      lo = mid;
  }
  return lo;
}

OPTK_HD double interp_at(const double *xp, const double *fp, int count, double x, int j) {
  if (x == xp[j]) return fp[j];
  double slope = (fp[j + 1] - fp[j]) / (xp[j + 1] - xp[j]);
  double r = slope * (x - xp[j]) + fp[j];
  if (r != r) {
    r = slope * (x - xp[j + 1]) + fp[j + 1];
    if (r != r && fp[j] == fp[j + 1]) r = fp[j];
  }
  return r;
}

OPTK_HD double interp1(const double *xp, const double *fp, int count, double x) {
  if (!(x >= xp[0])) return (x != x) ? x : fp[0];  // left clamp (NaN passes)
  if (x >= xp[count - 1]) return fp[count - 1];
  int j = interp_bracket(xp, count, x);
  return interp_at(xp, fp, count, x, j);
}

// the same rules over float32 tables
OPTK_HD float interp1f(const float *xp, const float *fp, int count, float x) {
  if (!(x >= xp[0])) return (x != x) ? x : fp[0];
  if (x >= xp[count - 1]) return fp[count - 1];
  int lo = 0, hi = count - 1;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (x < xp[mid])
      hi = mid;
    else
      lo = mid;
  }
  if (x == xp[lo]) return fp[lo];
  return (fp[lo + 1] - fp[lo]) / (xp[lo + 1] - xp[lo]) * (x - xp[lo]) + fp[lo];
}

// Index grid over one chemical's wavelength knots.  Measured n/k tables
// are often log-spaced over many octaves (0.1 A .. 10^6 A), so the cells are
// uniform in the bit pattern of the float32 wavelength -- a piecewise-linear
// log2: cell = (bits(x) - base) >> shift, 2^(23 - shift) cells per octave.
// lo[g] is the knot bracket at cell g's left edge.  A lookup starts from its
// cell's entry and steps to the bracket w[j] <= x < w[j + 1]; the bracket is
// unique, so the result is the binary search's whatever the cell (float32
// rounding of a float64 x included), in one load and a step or two instead
// of log2(count) dependent loads.  cells == 0: no grid (binary search).
struct ChemGrid {
  int32_t base, shift;    // float32 bits of w[0], cell width in bits (log2)
  int32_t cells, offset;  // cell count, offset into the lo array
};

OPTK_HD int grid_cell(const ChemGrid &g, float x) {
  int32_t b;
  memcpy(&b, &x, sizeof(b));
  int32_t c = b < g.base ? 0 : (b - g.base) >> g.shift;
  return c >= g.cells ? g.cells - 1 : c;
}

template <typename T>
OPTK_HD int grid_bracket(const ChemGrid &g, const int32_t *lo_tab, const T *w, int cnt, T x) {
  int j = lo_tab[g.offset + grid_cell(g, (float)x)];
  while (j > 0 && w[j] > x) --j;
  while (j < cnt - 2 && w[j + 1] <= x) ++j;
  return j;
}

// host: the grids of every chemical, about 8 cells per knot (at most 16384);
// a chemical whose first knot is not a positive normal float gets none
inline void build_chem_grids(const optk_chemical *chems, int num_chemicals, const double *tables,
                             ChemGrid *grids, int32_t *lo_tab, int64_t *lo_count) {
  int64_t used = 0;
  for (int c = 0; c < num_chemicals; ++c) {
    const double *w = tables + chems[c].offset;
    const int cnt = chems[c].count;
    ChemGrid g = {0, 0, 0, (int32_t)used};
    const float f0 = (float)w[0], f1 = (float)w[cnt - 1];
    if (cnt >= 2 && f0 >= 1.17549435e-38f && f1 < 3.0e38f && f1 >= f0) {
      int32_t b0, b1;
      memcpy(&b0, &f0, sizeof(b0));
      memcpy(&b1, &f1, sizeof(b1));
      int64_t want = 8 * (int64_t)cnt;
      want = want < 16 ? 16 : (want > 16384 ? 16384 : want);
      int shift = 0;
      while (shift < 30 && (((int64_t)(b1 - b0) >> shift) + 1) > want) ++shift;
      g.base = b0;
      g.shift = shift;
      g.cells = (int32_t)(((int64_t)(b1 - b0) >> shift) + 1);
      if (lo_tab) {
        int j = 0;
        for (int k = 0; k < g.cells; ++k) {
          const int32_t eb = b0 + (int32_t)((int64_t)k << shift);
          float edge;
          memcpy(&edge, &eb, sizeof(edge));
          while (j < cnt - 2 && w[j + 1] <= (double)edge) ++j;
          lo_tab[used + k] = j;
        }
      }
    }
    if (grids) grids[c] = g;
    used += g.cells;
  }
  if (lo_count) *lo_count = used;
}

// n + i k of a chemical at wavelength (nm); the table is in Angstrom.
// numpy.interp of a complex fp interpolates both parts with one reciprocal
// of the knot spacing (numpy's complex interp loop); the same is done here.
// With `grid`, the bracket comes from the chemical's index grid.
OPTK_HD C<double> chemical_n(const optk_chemical &ch, const double *tables, double wavelength_nm,
                             const ChemGrid *grid = nullptr, const int32_t *lo_tab = nullptr) {
  const double *w = tables + ch.offset;
  const double *n = w + ch.count;
  const double *k = n + ch.count;
  double x = wavelength_nm * 10.0;
  int cnt = ch.count;
  if (!(x >= w[0])) {
    if (x != x) return {x, x};
    return {n[0], k[0]};
  }
  if (x >= w[cnt - 1]) return {n[cnt - 1], k[cnt - 1]};
  int j = grid && grid->cells ? grid_bracket(*grid, lo_tab, w, cnt, x) : interp_bracket(w, cnt, x);
  if (x == w[j]) return {n[j], k[j]};
  double inv_dx = 1.0 / (w[j + 1] - w[j]);
  double sre = (n[j + 1] - n[j]) * inv_dx;
  double sim = (k[j + 1] - k[j]) * inv_dx;
  double re = sre * (x - w[j]) + n[j];
  double im = sim * (x - w[j]) + k[j];
  if (re != re) {
    re = sre * (x - w[j + 1]) + n[j + 1];
    if (re != re && n[j] == n[j + 1]) re = n[j];
  }
  if (im != im) {
    im = sim * (x - w[j + 1]) + k[j + 1];
    if (im != im && k[j] == k[j + 1]) im = k[j];
  }
  return {re, im};
}

// ---------------------------------------------------------------------------
// multilayer transfer matrices (optika/materials/_layers.py, matrices.py,
// _multilayers.py, profiles.py)
// ---------------------------------------------------------------------------

struct StackView {
  const ChemGrid *grids;   // per chemical, or null (binary search)
  const int32_t *grid_lo;
  const optk_layer *layers;
  const optk_stack *stacks;
  const optk_chemical *chemicals;
  const double *tables;
};

// snells_law_scalar (optika/materials/_snells_law.py:13-38)
template <typename T> OPTK_HD C<T> snell_scalar(C<T> cos_i, C<T> n1, C<T> n2) {
  C<T> one = {T(1), T(0)};
  C<T> sin_i = csqrt(csub(one, cmul(cos_i, cos_i)));
  C<T> sin_t = cdiv(cmul(n1, sin_i), n2);
  return csqrt(csub(one, cmul(sin_t, sin_t)));
}

// interface roughness factor (profiles.py:103-126 and each profile)
template <typename T>
OPTK_HD T profile_factor(int kind, T width, T wavelength, C<T> direction, C<T> n) {
  // k = -2 pi n direction / wavelength ; s = Re(-2 k)
  C<T> k = cdiv(cmul(cscale(T(-2) * T(M_PI), n), direction), C<T>{wavelength, T(0)});
  T s = T(-2) * k.re;
  switch (kind) {
  case OPTK_PROFILE_ERF: {
    T sw = s * width;
    return exp(-(sw * sw) / T(2));
  }
  case OPTK_PROFILE_EXPONENTIAL: {
    T sw = s * width;
    return T(1) / (T(1) + (sw * sw) / T(2));
  }
  case OPTK_PROFILE_LINEAR: {
    T x = T(1.7320508075688772) * width * s;
    return sin(x) / x;
  }
  case OPTK_PROFILE_SINUSOIDAL: {
    T a = T(M_PI) / (T(M_PI) * T(M_PI) - T(8));
    T x = a * width * s;
    T x1 = x - T(M_PI) / T(2);
    T x2 = x + T(M_PI) / T(2);
    return T(M_PI) * (sin(x1) / x1 + sin(x2) / x2) / T(4);
  }
  default:
    return T(1);
  }
}

template <typename T> struct LayerState {
  C<T> n;          // index of the medium before the next layer
  C<T> direction;  // cosine in that medium
  bool where;      // transfer matrices still valid
};

// transfer matrices of both polarizations (s, p)
template <typename T> struct M2Pair {
  M2<T> s, p;
};

// matrices.refraction (matrices.py:127-161) for one polarization, from the
// incident medium (n_i, cos_i) into (n_j, cos_j)
template <typename T>
OPTK_HD M2<T> refraction_matrix(const optk_layer &layer, T wavelength, bool pol_s, C<T> n_i,
                                C<T> cos_i, C<T> n_j, C<T> cos_j) {
  C<T> one = {T(1), T(0)};
  C<T> di = pol_s ? cos_i : cconj(cos_i);
  C<T> dj = pol_s ? cos_j : cconj(cos_j);
  C<T> imp_i = pol_s ? n_i : cdiv(one, n_i);
  C<T> imp_j = pol_s ? n_j : cdiv(one, n_j);
  C<T> qi = cmul(di, imp_i);
  C<T> qj = cmul(dj, imp_j);
  C<T> a = cadd(qi, qj);
  C<T> r = cdiv(csub(qi, qj), a);
  C<T> t = cdiv(cscale(T(2), qi), a);
  if (layer.profile_kind != OPTK_PROFILE_NONE) {
    T f = profile_factor<T>(layer.profile_kind, T(layer.profile_width), wavelength, di, n_i);
    r = cscale(f, r);  // r * factor (real factor promoted to complex)
  }
  C<T> inv_t = cdiv(one, t);
  C<T> r_t = cdiv(r, t);
  return {inv_t, r_t, r_t, inv_t};
}

// Layer.transfer (optika/materials/_layers.py:229-277) for both
// polarizations at once.  Everything but the refraction matrix (the index,
// the refracted cosine, the propagation phases and the validity guard)
// depends on the medium only, so it is evaluated once; each polarization's
// matrix keeps the reference's operation order, so the result is the same
// as two separate per-polarization passes.  `n_int` is the layer's index
// at this wavelength (chemical_n).
template <typename T>
OPTK_HD M2Pair<T> layer_transfer2(const optk_layer &layer, C<T> n_int, T wavelength,
                                  LayerState<T> &st, T thickness) {
  C<T> d_int = snell_scalar(st.direction, st.n, n_int);
  M2Pair<T> refr;
  if (st.where) {
    refr.s = refraction_matrix<T>(layer, wavelength, true, st.n, st.direction, n_int, d_int);
    refr.p = refraction_matrix<T>(layer, wavelength, false, st.n, st.direction, n_int, d_int);
  } else {
    refr.s = midentity<T>();
    refr.p = midentity<T>();
  }
  // matrices.propagation (matrices.py:241-246):
  // beta = 2 pi h n cos / wavelength ; U = diag(exp(-i beta), exp(i beta))
  C<T> beta =
      cdiv(cmul(cscale(T(2.0 * M_PI) * thickness, n_int), d_int), C<T>{wavelength, T(0)});
  // -1j * beta = (beta.im, -beta.re) ; +1j * beta = (-beta.im, beta.re)
  C<T> u00 = cexp(C<T>{beta.im, -beta.re});
  C<T> u11 = cexp(C<T>{-beta.im, beta.re});
  bool where_prop = cabs(u00) < T(1e10);
  st.where = st.where && where_prop;
  M2<T> zero_prop = {u00, {T(0), T(0)}, {T(0), T(0)}, u11};
  M2Pair<T> out;
  out.s = st.where ? mmul(refr.s, zero_prop) : refr.s;
  out.p = st.where ? mmul(refr.p, zero_prop) : refr.p;
  st.n = n_int;
  st.direction = d_int;
  return out;
}

template <typename T> OPTK_HD M2<T> matrix_power(M2<T> m, int exponent) {
  // binary powering, numpy.linalg.matrix_power's order of products
  if (exponent <= 0) return midentity<T>();
  M2<T> z = m, result = m;
  bool have_z = false, have_result = false;
  int e = exponent;
  while (e > 0) {
    if (have_z) z = mmul(z, z);
    have_z = true;
    int bit = e & 1;
    e >>= 1;
    if (bit) {
      result = have_result ? mmul(result, z) : z;
      have_result = true;
    }
  }
  return result;
}

template <typename T>
OPTK_HD C<T> layer_index(const optk_layer &L, const StackView &sv, T wavelength) {
  if (L.chemical < 0) return {T(1), T(0)};
  return to_c<T>(chemical_n(sv.chemicals[L.chemical], sv.tables, (double)wavelength,
                            sv.grids ? sv.grids + L.chemical : nullptr, sv.grid_lo));
}

// LayerSequence.transfer over layers first .. first + count - 1, both
// polarizations; `n_cache` holds the indices of the first 4 layers when
// `cached` (a periodic block's second pass reuses its first pass's lookups)
template <typename T>
OPTK_HD M2Pair<T> sequence_transfer2(const StackView &sv, int first, int count, T wavelength,
                                     LayerState<T> &st, C<T> n_cache[4], bool cached) {
  M2Pair<T> result = {midentity<T>(), midentity<T>()};
  for (int i = 0; i < count; ++i) {
    const optk_layer &L = sv.layers[first + i];
    C<T> n_int;
    if (i < 4 && cached) {
      n_int = n_cache[0];
      if (i == 1) n_int = n_cache[1];
      if (i == 2) n_int = n_cache[2];
      if (i == 3) n_int = n_cache[3];
    } else {
      n_int = layer_index<T>(L, sv, wavelength);
      if (i == 0) n_cache[0] = n_int;
      if (i == 1) n_cache[1] = n_int;
      if (i == 2) n_cache[2] = n_int;
      if (i == 3) n_cache[3] = n_int;
    }
    M2Pair<T> x = layer_transfer2<T>(L, n_int, wavelength, st, T(L.thickness));
    result.s = mmul(result.s, x.s);
    result.p = mmul(result.p, x.p);
  }
  return result;
}

template <typename T> struct MultilayerOut {
  C<T> r[2];  // s, p
  C<T> t[2];
  bool where[2];
};

// multilayer_coefficients (optika/materials/_multilayers.py:186-237)
template <typename T>
OPTK_HD MultilayerOut<T> multilayer_rt(const StackView &sv, int stack_index, T wavelength,
                                       C<T> direction, C<T> n_ambient) {
  const optk_stack &stk = sv.stacks[stack_index];
  MultilayerOut<T> out;
  LayerState<T> st = {n_ambient, direction, true};
  M2Pair<T> m = {midentity<T>(), midentity<T>()};
  bool first_block = true;
  for (int b = 0; b < stk.num_blocks; ++b) {
    const optk_block &blk = stk.blocks[b];
    C<T> n_cache[4];
    M2Pair<T> x;
    if (blk.kind == OPTK_BLOCK_PERIODIC) {
      M2Pair<T> start =
          sequence_transfer2<T>(sv, blk.first_layer, blk.num_layers, wavelength, st, n_cache, false);
      M2Pair<T> periodic =
          sequence_transfer2<T>(sv, blk.first_layer, blk.num_layers, wavelength, st, n_cache, true);
      x.s = mmul(start.s, matrix_power(periodic.s, blk.num_periods - 1));
      x.p = mmul(start.p, matrix_power(periodic.p, blk.num_periods - 1));
    } else {
      x = sequence_transfer2<T>(sv, blk.first_layer, blk.num_layers, wavelength, st, n_cache, false);
    }
    // a single top-level block is returned by its own transfer(); several
    // blocks are folded left from the identity (LayerSequence.transfer)
    if (first_block && stk.num_blocks == 1) {
      m = x;
    } else {
      m.s = mmul(m.s, x.s);
      m.p = mmul(m.p, x.p);
    }
    first_block = false;
  }
  // substrate (thickness forced to zero)
  optk_layer sub = {-1, OPTK_PROFILE_NONE, 0.0, 0.0};
  if (stk.substrate_layer >= 0) sub = sv.layers[stk.substrate_layer];
  M2Pair<T> ms = layer_transfer2<T>(sub, layer_index<T>(sub, sv, wavelength), wavelength, st, T(0));
  M2<T> mt[2] = {mmul(m.s, ms.s), mmul(m.p, ms.p)};
  for (int p = 0; p < 2; ++p) {
    out.r[p] = cdiv(mt[p].c, mt[p].a);
    C<T> t = cdiv(C<T>{T(1), T(0)}, mt[p].a);
    if (!st.where) t = {T(0), T(0)};
    out.t[p] = t;
    out.where[p] = st.where;
  }
  return out;
}

template <typename Real> struct MultilayerRT {
  Real R[2], T[2];
};

// multilayer_efficiency (optika/materials/_multilayers.py:485-532)
template <typename T>
OPTK_HD MultilayerRT<T> multilayer_efficiency(const StackView &sv, int stack_index, T wavelength,
                                              C<T> direction, C<T> n_ambient,
                                              bool want_transmission,
                                              MultilayerOut<T> *coeffs = nullptr) {
  MultilayerOut<T> co = multilayer_rt<T>(sv, stack_index, wavelength, direction, n_ambient);
  if (coeffs) *coeffs = co;
  MultilayerRT<T> out;
  for (int p = 0; p < 2; ++p) {
    T a = cabs(co.r[p]);
    out.R[p] = a * a;
    out.T[p] = T(0);
  }
  if (want_transmission) {
    const optk_stack &stk = sv.stacks[stack_index];
    C<T> n_sub = {T(1), T(0)};
    if (stk.substrate_layer >= 0) {
      const optk_layer &sub = sv.layers[stk.substrate_layer];
      n_sub = layer_index<T>(sub, sv, wavelength);
    }
    C<T> d_sub = snell_scalar(direction, n_ambient, n_sub);
    C<T> one = {T(1), T(0)};
    C<T> q_amb_s = cmul(direction, n_ambient);
    C<T> q_amb_p = cmul(cconj(direction), cdiv(one, n_ambient));
    C<T> q_sub_s = cmul(d_sub, n_sub);
    C<T> q_sub_p = cmul(cconj(d_sub), cdiv(one, n_sub));
    T ts = cabs(co.t[0]);
    T tp = cabs(co.t[1]);
    out.T[0] = (ts * ts) * cdiv(q_sub_s, q_amb_s).re;
    out.T[1] = (tp * tp) * cdiv(q_sub_p, q_amb_p).re;
  }
  return out;
}

}  // namespace optk
