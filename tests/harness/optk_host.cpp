// optk_host.cpp -- TEST HARNESS ONLY: the trace kernels' per-ray code
// (optika_amd/csrc/optk_surface.h) compiled for the host with g++, so the
// device arithmetic can be compared with the numpy oracle on a machine
// without a GPU.  Never linked into or loaded by optika_amd.
#include <cstdint>
#include <vector>

#include "../../optika_amd/csrc/optk_surface.h"

using namespace optk;

namespace {
SystemView view_of(const optk_system_desc *d) {
  SystemView v;
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
