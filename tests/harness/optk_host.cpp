// optk_host.cpp -- TEST HARNESS ONLY: the trace kernels' per-ray code
// (optika_amd/csrc/optk_surface.h) compiled for the host with g++, so the
// device arithmetic can be compared with the numpy oracle on a machine
// without a GPU.  Never linked into or loaded by optika_amd.
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../optika_amd/csrc/optk_fast.h"
#include "../../optika_amd/csrc/optk_random.h"

using namespace optk;

namespace {
// n/k lookup grids of the last system viewed (host copies)
std::vector<ChemGrid> g_grids;
std::vector<int32_t> g_grid_lo;

SystemView view_of(const optk_system_desc *d) {
  int64_t count = 0;
  build_chem_grids(d->chemicals, d->num_chemicals, d->tables, nullptr, nullptr, &count);
  g_grids.assign(d->num_chemicals > 0 ? d->num_chemicals : 1, ChemGrid());
  g_grid_lo.assign(count > 0 ? count : 1, 0);
  build_chem_grids(d->chemicals, d->num_chemicals, d->tables, g_grids.data(), g_grid_lo.data(), nullptr);
  SystemView v;
  v.sv.grids = d->num_chemicals > 0 ? g_grids.data() : nullptr;
  v.sv.grid_lo = g_grid_lo.data();
  v.surfaces = d->surfaces;
  v.num_surfaces = d->num_surfaces;
  v.num_paths = d->num_paths;
  v.sv.layers = d->layers;
  v.sv.stacks = d->stacks;
  v.sv.chemicals = d->chemicals;
  v.sv.tables = d->tables;
  return v;
}
}  // namespace

std::vector<optk_surface> derived(const optk_system_desc *d) {
  std::vector<optk_surface> v(d->surfaces, d->surfaces + d->num_surfaces);
  for (auto &s : v) derive_constants(s);
  return v;
}

extern "C" {

// Philox4x32-10 of one counter under one key (known-answer tests)
void host_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  uint32_t k[2] = {key[0], key[1]};
  for (int r = 0; r < 10; ++r) {
    philox_round(c, k);
    k[0] += 0x9E3779B9u;
    k[1] += 0xBB67AE85u;
  }
  for (int j = 0; j < 4; ++j) out[j] = c[j];
}

// the exposure's draws for pixels 0 .. n-1: Poisson of mean[i] plus
// read_noise times a Gaussian (the kernel's per-pixel code)
void host_expose(const double *mean, int64_t n, uint64_t seed, double read_noise, double *out) {
  for (int64_t i = 0; i < n; ++i) {
    UniformStream rng;
    rng.pixel = (uint64_t)i;
    rng.seed = seed;
    double e = poisson_draw(mean[i], rng);
    if (read_noise > 0.0) e += read_noise * gaussian_draw(rng);
    out[i] = e;
  }
}

// Knot brackets of n wavelengths (Angstrom) in one ascending table through
// its index grid: j[i] (w[j] <= x < w[j + 1], clamped to 0 .. count - 2) and
// steps[i], the knots walked from the grid cell's entry.  Returns the
// grid's cell count (0: no grid).
int host_grid_bracket(const double *w, int count, const double *x, int n, int32_t *j,
                      int32_t *steps) {
  optk_chemical ch = {0, count};
  ChemGrid g;
  int64_t cells = 0;
  build_chem_grids(&ch, 1, w, nullptr, nullptr, &cells);
  std::vector<int32_t> lo(cells > 0 ? cells : 1, 0);
  build_chem_grids(&ch, 1, w, &g, lo.data(), nullptr);
  for (int i = 0; i < n; ++i) {
    if (!g.cells) {
      j[i] = -1;
      steps[i] = 0;
      continue;
    }
    const int start = lo[grid_cell(g, (float)x[i])];
    j[i] = grid_bracket(g, lo.data(), w, count, x[i]);
    steps[i] = j[i] > start ? j[i] - start : start - j[i];
  }
  return g.cells;
}

// Trace n grid rays; out[0..9] = px py pz dx dy dz wavelength intensity
// attenuation index (num_paths * n each, path-major), mask = unvignetted.
int host_trace_grid(const optk_system_desc *d, const double *fx, const double *fy,
                    const double *ux, const double *uy, const double *wl, const double *inten,
                    int64_t n, int object_at_infinity, int post, double **out,
                    uint8_t *mask) {
  std::vector<optk_surface> surf = derived(d);
  TraceSpec P;
  P.sys = view_of(d);
  P.sys.surfaces = surf.data();
  P.pre[0] = d->pre[0];
  P.pre[1] = d->pre[1];
  P.pre_is_inverse[0] = d->pre_is_inverse[0];
  P.pre_is_inverse[1] = d->pre_is_inverse[1];
  P.post = d->post;
  P.apply_post = post;
  for (int64_t i = 0; i < n; ++i) {
    Ray<double> r0 = grid_ray(fx[i], fy[i], ux[i], uy[i], wl[i], inten[i], object_at_infinity);
    apply_pre(P, r0);
    for (int path = 0; path < d->num_paths; ++path) {
      Ray<double> r = r0;
      for (int k = 0; k < d->num_surfaces; ++k) propagate_surface(surf[k], P.sys, path, r);
      apply_post(P, r);
      int64_t j = (int64_t)path * n + i;
      double v[10] = {r.px, r.py, r.pz, r.dx, r.dy, r.dz, r.wavelength, r.intensity,
                      r.attenuation, r.index};
      for (int c = 0; c < 10; ++c) out[c][j] = v[c];
      mask[j] = r.unvignetted ? 1 : 0;
    }
  }
  return 0;
}

// The float32 form (optk_fast.h) on the host: float32 grid in, float32 rays
// out (same layout as host_trace_grid).
int host_trace_grid_f32(const optk_system_desc *d, const float *fx, const float *fy,
                        const float *ux, const float *uy, const float *wl, const float *inten,
                        int64_t n, int object_at_infinity, int post, float **out,
                        uint8_t *mask) {
  std::vector<optk_surface> surf = derived(d);
  std::vector<FastSurface> fast(d->num_surfaces);
  FastSpec spec;
  std::memset(&spec, 0, sizeof(spec));
  fast_build(surf.data(), d->num_surfaces, d->pre, d->pre_is_inverse, d->post, fast.data(), spec);
  std::vector<float> tables32(d->tables, d->tables + d->num_tables);
  FastView v;
  v.surfaces = surf.data();
  v.fast = fast.data();
  v.sv = view_of(d).sv;
  v.tables32 = tables32.data();
  for (int64_t i = 0; i < n; ++i) {
    FRay r0 = fast_grid_ray(fx[i], fy[i], ux[i], uy[i], wl[i], inten[i], object_at_infinity);
    // as the grid kernels: a pass-through surface 0 is elided
    const int k0 = spec.skip_first ? 1 : 0;
    if (spec.skip_first) {
      const bool parallel = r0.dz == 0.0;
      fast_xf(spec.first_skip, r0);
      if (parallel) {
        r0.px = r0.py = r0.pz = NAN;
        r0.intensity = NAN;
      }
    } else {
      fast_xf(spec.first_hop[1], r0);
    }
    for (int path = 0; path < d->num_paths; ++path) {
      FRay r = r0;
      for (int k = k0; k < d->num_surfaces; ++k) {
        if (k > k0) fast_xf(fast[k].hop, r);
        fast_propagate<FEAT_ALL>(fast[k], surf[k], v, path, r);
      }
      fast_xf(spec.last[post ? 1 : 0], r);
      int64_t j = (int64_t)path * n + i;
      float vals[10] = {(float)r.px, (float)r.py, (float)r.pz, (float)r.dx, (float)r.dy,
                        (float)r.dz, r.wavelength, r.intensity, r.attenuation, r.index};
      for (int c = 0; c < 10; ++c) out[c][j] = vals[c];
      mask[j] = r.unvignetted ? 1 : 0;
    }
  }
  return 0;
}

// the float64 kernels' small-angle sin and cos (direction from field angles)
void host_small_angle_sincos(const double *x, double *s, double *c, int64_t n) {
  for (int64_t i = 0; i < n; ++i) small_angle_sincos(x[i], s[i], c[i]);
}

// 1 when the float32 form's grid kernels elide surface 0 of the system
// (FastSpec::skip_first), else 0.
int host_fast_skip_first(const optk_system_desc *d) {
  std::vector<optk_surface> surf = derived(d);
  std::vector<FastSurface> fast(d->num_surfaces);
  FastSpec spec;
  std::memset(&spec, 0, sizeof(spec));
  fast_build(surf.data(), d->num_surfaces, d->pre, d->pre_is_inverse, d->post, fast.data(), spec);
  return spec.skip_first;
}

// The float32 form over a surface sequence: float32 rays in (system frame),
// float32 rays out.  in[0..9] / out[0..9] = px py pz dx dy dz wavelength
// intensity attenuation index.
int host_trace_rays_f32(const optk_system_desc *d, const float *const *in, int64_t n,
                        float **out, uint8_t *mask) {
  std::vector<optk_surface> surf = derived(d);
  std::vector<FastSurface> fast(d->num_surfaces);
  FastSpec spec;
  std::memset(&spec, 0, sizeof(spec));
  fast_build(surf.data(), d->num_surfaces, d->pre, d->pre_is_inverse, d->post, fast.data(), spec);
  std::vector<float> tables32(d->tables, d->tables + d->num_tables);
  FastView v;
  v.surfaces = surf.data();
  v.fast = fast.data();
  v.sv = view_of(d).sv;
  v.tables32 = tables32.data();
  for (int64_t i = 0; i < n; ++i) {
    FRay r0;
    r0.px = in[0][i];
    r0.py = in[1][i];
    r0.pz = in[2][i];
    r0.dx = in[3][i];
    r0.dy = in[4][i];
    r0.dz = in[5][i];
    r0.wavelength = in[6][i];
    r0.intensity = in[7][i];
    r0.attenuation = in[8][i];
    r0.index = in[9][i];
    r0.unvignetted = true;
    xf_apply_pos(spec.first_hop[0], r0.px, r0.py, r0.pz);
    xf_apply_dir(spec.first_hop[0], r0.dx, r0.dy, r0.dz);
    for (int path = 0; path < d->num_paths; ++path) {
      FRay r = r0;
      for (int k = 0; k < d->num_surfaces; ++k) {
        if (k > 0) {
          xf_apply_pos(fast[k].hop, r.px, r.py, r.pz);
          xf_apply_dir(fast[k].hop, r.dx, r.dy, r.dz);
        }
        fast_propagate<FEAT_ALL>(fast[k], surf[k], v, path, r);
      }
      xf_apply_pos(spec.last[0], r.px, r.py, r.pz);
      xf_apply_dir(spec.last[0], r.dx, r.dy, r.dz);
      int64_t j = (int64_t)path * n + i;
      float vals[10] = {(float)r.px, (float)r.py, (float)r.pz, (float)r.dx, (float)r.dy,
                        (float)r.dz, r.wavelength, r.intensity, r.attenuation, r.index};
      for (int c = 0; c < 10; ++c) out[c][j] = vals[c];
      mask[j] = r.unvignetted ? 1 : 0;
    }
  }
  return 0;
}

// Multilayer reflectivity/transmissivity of stack `stack` (s, p).
int host_multilayer(const optk_system_desc *d, int stack, const double *wl, const double *cos_re,
                    const double *cos_im, const double *n_re, const double *n_im, int64_t n,
                    double *Rs, double *Rp, double *Ts, double *Tp) {
  SystemView v = view_of(d);
  for (int64_t i = 0; i < n; ++i) {
    MultilayerRT<double> rt = multilayer_efficiency<double>(v.sv, stack, wl[i], C<double>{cos_re[i], cos_im[i]},
                                            C<double>{n_re[i], n_im[i]}, true);
    Rs[i] = rt.R[0];
    Rp[i] = rt.R[1];
    Ts[i] = rt.T[0];
    Tp[i] = rt.T[1];
  }
  return 0;
}
}
