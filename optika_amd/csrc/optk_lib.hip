// optk_lib.hip -- gfx950 kernels and the C ABI of liboptk.so (float64 trace,
// per-operator evaluation, sensor binning).  See include/optk.h.
//
// Compiled with -ffp-contract=off: the float64 trace keeps the reference's
// rounding sequence (no fused multiply-adds).

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "optk_batch.h"
#include "optk_random.h"
#include "optk_signatures.h"

using namespace optk;

namespace {

thread_local std::string g_last_error;

}  // namespace

namespace optk {
int32_t fail(int32_t code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

int64_t resident_grid(const void *kernel, int block, int64_t n) {
  static std::mutex mu;
  static std::map<std::pair<const void *, int>, int64_t> cache;
  const int64_t need = (n + block - 1) / block;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return need < 4096 ? need : 4096;
  int64_t resident = 0;
  {
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find({kernel, dev});
    if (it != cache.end()) resident = it->second;
  }
  if (resident == 0) {
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) == hipSuccess &&
        cus > 0 && per_cu > 0)
      resident = (int64_t)cus * per_cu * 4;
    else
      resident = 4096;
    std::lock_guard<std::mutex> lock(mu);
    cache[{kernel, dev}] = resident;
  }
  return need < resident ? need : resident;
}
}  // namespace optk

#define OPTK_HIP_CHECK(expr)                                                         \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail(OPTK_EDEVICE, std::string(#expr " failed: ") + hipGetErrorString(e_)); \
  } while (0)

static int32_t features_of(const optk_system_desc *d) {
  int32_t f = 0;
  for (int i = 0; i < d->num_surfaces; ++i) {
    const optk_surface &s = d->surfaces[i];
    if (s.sag_kind == OPTK_SAG_TOROID) f |= FEAT_TOROID;
    if (s.rulings_kind != OPTK_RULINGS_NONE) f |= FEAT_RULINGS;
    if (s.material_kind == OPTK_MATERIAL_MULTILAYER_MIRROR ||
        s.material_kind == OPTK_MATERIAL_MULTILAYER_FILM ||
        s.material_kind == OPTK_MATERIAL_SILICON_SENSOR)
      f |= FEAT_MULTILAYER;
  }
  return f;
}

// kernel variant that covers a feature set: geometric only, rulings and
// toroids without multilayers, or everything
static int32_t variant_of(int32_t features) {
  if (features == 0) return 0;
  if (!(features & FEAT_MULTILAYER)) return FEAT_TOROID | FEAT_RULINGS;
  return FEAT_ALL;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// Per-launch arguments.  The system (surfaces, tables, transformations) is
// device-resident; a grid-stride loop traces the rays, reading the records
// with scalar loads (the batched and stop-solve kernels stage them in LDS).
struct TraceArgs {
  const TraceSpec *spec;         // device copy
  const optk_surface *surfaces;  // device copy
  int32_t num_surfaces;
  int32_t apply_post;
  int64_t n;
  RaysOut out;
};

__device__ __forceinline__ double ld(const double *p, int64_t i, double dflt) {
  return p ? p[i] : dflt;
}

// a streaming store: the trace writes each output element once and never
// reads it back, so the store is marked nontemporal and does not keep the
// line in the caches (the Newtonian sensor-frame kernels measured 8 %
// faster in float32 and 2-3 % in float64 with it; nontemporal loads of the
// inputs gained nothing)
template <typename T> __device__ __forceinline__ void put(T &dst, T v) { __builtin_nontemporal_store(v, &dst); }

__device__ __forceinline__ void store(const RaysOut &o, int64_t j, const Ray<double> &r) {
  if (o.p[0]) put<double>(o.p[0][j], r.px);
  if (o.p[1]) put<double>(o.p[1][j], r.py);
  if (o.p[2]) put<double>(o.p[2][j], r.pz);
  if (o.d[0]) put<double>(o.d[0][j], r.dx);
  if (o.d[1]) put<double>(o.d[1][j], r.dy);
  if (o.d[2]) put<double>(o.d[2][j], r.dz);
  if (o.wavelength) put<double>(o.wavelength[j], r.wavelength);
  if (o.intensity) put<double>(o.intensity[j], r.intensity);
  if (o.attenuation) put<double>(o.attenuation[j], r.attenuation);
  if (o.index) put<double>(o.index[j], r.index);
  if (o.unvignetted) put<uint8_t>(o.unvignetted[j], r.unvignetted ? 1 : 0);
}

__device__ __forceinline__ Ray<double> load_ray(const RaysIn &in, int64_t i) {
  Ray<double> r;
  r.px = in.p[0][i];
  r.py = in.p[1][i];
  r.pz = in.p[2][i];
  r.dx = in.d[0][i];
  r.dy = in.d[1][i];
  r.dz = in.d[2][i];
  r.wavelength = ld(in.wavelength, i, in.wavelength_default);
  r.intensity = ld(in.intensity, i, in.intensity_default);
  r.attenuation = ld(in.attenuation, i, in.attenuation_default);
  r.index = ld(in.index, i, in.index_default);
  r.unvignetted = in.unvignetted ? (in.unvignetted[i] != 0) : (in.unvignetted_default != 0);
  return r;
}

__device__ __forceinline__ Ray<double> load_ray(const GridIn &g, int64_t i) {
  return grid_ray(ld(g.field[0], i, g.field_default[0]), ld(g.field[1], i, g.field_default[1]),
                  ld(g.pupil[0], i, g.pupil_default[0]), ld(g.pupil[1], i, g.pupil_default[1]),
                  ld(g.wavelength, i, g.wavelength_default), ld(g.intensity, i, g.intensity_default),
                  g.object_at_infinity);
}

// bytes of LDS a batched or stop-solve launch stages: the spec, then the
// surface records
static size_t staged_bytes(int num_surfaces) {
  return sizeof(TraceSpec) + sizeof(optk_surface) * (size_t)num_surfaces;
}

// occupancy targets: the geometric kernels keep <= 128 VGPRs (4 waves per
// SIMD) to hide HBM latency; the multilayer kernels are VALU-bound
template <int F> struct Waves {
  static constexpr int value = (F & FEAT_MULTILAYER) ? 2 : 4;
};
// the specialised kernels: 4 waves also for the multilayer layout (128
// VGPRs, more of the multilayer's temporaries spilled), measured 2.5 %
// faster on the spectrograph than 2 waves with 256 VGPRs
template <int F> struct SpecWaves {
  static constexpr int value = 4;
};

// I/O layouts of the kernels.  IO_ANY honours every null pointer of the
// descriptors (a value default in, the field not written out).  IO_SENSOR
// is the layout of the sensor-frame rayfunction: field, pupil and
// wavelength arrays in, the intensity one value; position, direction,
// intensity and the mask out, nothing else.  With the presence fixed at
// compile time the kernel keeps no null tests, and none of the pointers it
// does not use, live across its ray loop (they would otherwise crowd the
// scalar registers the system records are read into).
// IO_FULL: every output field is written (the ray entry point's usual
// layout); the inputs keep their null tests.
enum { IO_ANY = 0, IO_SENSOR = 1, IO_FULL = 2 };

static bool full_output(const RaysOut &o) {
  return o.p[0] && o.p[1] && o.p[2] && o.d[0] && o.d[1] && o.d[2] && o.wavelength && o.intensity &&
         o.attenuation && o.index && o.unvignetted;
}

static bool sensor_layout(const GridIn &g, const RaysOut &o) {
  return g.field[0] && g.field[1] && g.pupil[0] && g.pupil[1] && g.wavelength && !g.intensity &&
         o.p[0] && o.p[1] && o.p[2] && o.d[0] && o.d[1] && o.d[2] && o.intensity &&
         o.unvignetted && !o.wavelength && !o.attenuation && !o.index;
}

template <int IO> __device__ __forceinline__ Ray<double> load_io(const RaysIn &in, int64_t i) {
  return load_ray(in, i);
}

template <int IO> __device__ __forceinline__ Ray<double> load_io(const GridIn &g, int64_t i) {
  if (IO == IO_SENSOR)
    return grid_ray(g.field[0][i], g.field[1][i], g.pupil[0][i], g.pupil[1][i], g.wavelength[i],
                    g.intensity_default, g.object_at_infinity);
  return load_ray(g, i);
}

template <int IO> __device__ __forceinline__ void store_io(const RaysOut &o, int64_t j, const Ray<double> &r) {
  if (IO == IO_FULL) {
    put<double>(o.p[0][j], r.px);
    put<double>(o.p[1][j], r.py);
    put<double>(o.p[2][j], r.pz);
    put<double>(o.d[0][j], r.dx);
    put<double>(o.d[1][j], r.dy);
    put<double>(o.d[2][j], r.dz);
    put<double>(o.wavelength[j], r.wavelength);
    put<double>(o.intensity[j], r.intensity);
    put<double>(o.attenuation[j], r.attenuation);
    put<double>(o.index[j], r.index);
    put<uint8_t>(o.unvignetted[j], r.unvignetted ? 1 : 0);
    return;
  }
  if (IO == IO_SENSOR) {
    put<double>(o.p[0][j], r.px);
    put<double>(o.p[1][j], r.py);
    put<double>(o.p[2][j], r.pz);
    put<double>(o.d[0][j], r.dx);
    put<double>(o.d[1][j], r.dy);
    put<double>(o.d[2][j], r.dz);
    put<double>(o.intensity[j], r.intensity);
    put<uint8_t>(o.unvignetted[j], r.unvignetted ? 1 : 0);
    return;
  }
  store(o, j, r);
}

// p[i] through a 32-bit byte offset against the uniform base: global
// addressing with the base in scalar registers and one VGPR offset, no
// 64-bit address add per array.  Callers guarantee i * sizeof(T) < 2^32.
template <typename T> __device__ __forceinline__ T &at32(T *p, uint32_t i) {
  typedef typename std::conditional<std::is_const<T>::value, const char, char>::type Byte;
  return *reinterpret_cast<T *>(reinterpret_cast<Byte *>(p) + i * (uint32_t)sizeof(T));
}


// the sensor-frame layout's loads and stores at 32-bit offsets (the
// specialised grid kernels; other layouts take the 64-bit forms)
template <int IO> __device__ __forceinline__ Ray<double> load_io32(const GridIn &g, uint32_t i) {
  if (IO == IO_SENSOR)
    return grid_ray(at32(g.field[0], i), at32(g.field[1], i), at32(g.pupil[0], i), at32(g.pupil[1], i),
                    at32(g.wavelength, i), g.intensity_default, g.object_at_infinity);
  return load_io<IO>(g, (int64_t)i);
}
template <int IO> __device__ __forceinline__ Ray<double> load_io32(const RaysIn &in, uint32_t i) {
  return load_io<IO>(in, (int64_t)i);
}

template <int IO> __device__ __forceinline__ void store_io32(const RaysOut &o, uint32_t j, const Ray<double> &r) {
  if (IO == IO_SENSOR) {
    // the offset is formed here (an empty asm stops it being hoisted into
    // an earlier block as a 64-bit address, which the stores could then not
    // take as a scalar base plus a 32-bit offset)
    asm volatile("" : "+v"(j));
    put<double>(at32(o.p[0], j), r.px);
    put<double>(at32(o.p[1], j), r.py);
    put<double>(at32(o.p[2], j), r.pz);
    put<double>(at32(o.d[0], j), r.dx);
    put<double>(at32(o.d[1], j), r.dy);
    put<double>(at32(o.d[2], j), r.dz);
    put<double>(at32(o.intensity, j), r.intensity);
    put<uint8_t>(at32(o.unvignetted, j), r.unvignetted ? 1 : 0);
    return;
  }
  store_io<IO>(o, (int64_t)j, r);
}

// The system's records are read through the constant address space: every
// lane of a wave reads the same record at the same time, so the loads are
// scalar-cache reads whose values land in SGPRs, which the VALU
// instructions take as operands directly (no VGPRs hold record fields and
// branch conditions on them are uniform).  The arithmetic is unchanged.
typedef __attribute__((address_space(4))) const TraceSpec ConstTraceSpec;
typedef __attribute__((address_space(4))) const optk_surface ConstSurface;

template <int F, bool ACCUM, int IO, typename In>
__device__ __forceinline__ void trace_body(const TraceArgs &A, const In &in, const RaysOut &accum) {
  ConstTraceSpec *c_spec = (ConstTraceSpec *)A.spec;
  ConstSurface *c_surf = (ConstSurface *)A.surfaces;
  const TraceSpec &spec = *(const TraceSpec *)c_spec;
  SystemView sys = spec.sys;
  sys.surfaces = A.surfaces;
  // only rulings create diffraction-order paths: without them one path
  const int num_paths = (F & FEAT_RULINGS) ? sys.num_paths : 1;
  const int num_surfaces = A.num_surfaces;
  const int64_t n = A.n;
  // the launch block is TRACE_BLOCK threads: a compile-time stride keeps
  // it in scalar registers (blockDim.x would be a per-lane load)
  const int64_t stride = (int64_t)gridDim.x * TRACE_BLOCK;
  for (int64_t i = (int64_t)blockIdx.x * TRACE_BLOCK + threadIdx.x; i < n; i += stride) {
    Ray<double> r0 = load_io<IO>(in, i);
    apply_pre(spec, r0);
    for (int path = 0; path < num_paths; ++path) {
      Ray<double> r = r0;
      for (int k = 0; k < num_surfaces; ++k) {
        propagate_surface<F>(*(const optk_surface *)&c_surf[k], sys, path, r);
        if (ACCUM) store(accum, ((int64_t)k * num_paths + path) * n + i, r);
      }
      if (A.apply_post) {
        xf_inverse_ray(spec.post, r.px, r.py, r.pz, r.dx, r.dy, r.dz);
      }
      store_io<IO>(A.out, (int64_t)path * n + i, r);
    }
  }
}

template <int F, bool ACCUM, int IO = IO_ANY>
__global__ __launch_bounds__(TRACE_BLOCK) __attribute__((amdgpu_waves_per_eu(Waves<F>::value)))
void trace_rays_f64(const TraceArgs A, const RaysIn in, const RaysOut accum) {
  trace_body<F, ACCUM, IO>(A, in, accum);
}

template <int F, bool ACCUM, int IO = IO_ANY>
__global__ __launch_bounds__(TRACE_BLOCK) __attribute__((amdgpu_waves_per_eu(Waves<F>::value)))
void trace_grid_f64(const TraceArgs A, const GridIn g, const RaysOut accum) {
  trace_body<F, ACCUM, IO>(A, g, accum);
}

// launch the variant covering the system's features; accumulation always
// uses the all-features kernel
template <typename In>
static void launch_trace(void (*k0)(TraceArgs, In, RaysOut), void (*k3)(TraceArgs, In, RaysOut),
                         void (*k7)(TraceArgs, In, RaysOut), void (*k7a)(TraceArgs, In, RaysOut),
                         int features, bool accum, int64_t n, hipStream_t stream,
                         const TraceArgs &A, const In &in, const RaysOut &acc) {
  auto kern = accum ? k7a : (features == 0 ? k0 : (features == (FEAT_TOROID | FEAT_RULINGS) ? k3 : k7));
  // whole rounds of resident blocks; each block loops over its share
  const int64_t blocks = resident_grid((const void *)kern, TRACE_BLOCK, n);
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(TRACE_BLOCK), 0, stream, A, in, acc);
}


// Kernels specialised on a system's surface kinds (optk_signatures.h): the
// surfaces unrolled in order, each with its kinds fixed at compile time, one
// ray per thread.  (A grid-stride loop would keep every surface's record
// fields live in scalar registers across its iterations, more than the
// scalar register file holds; the compiler then spills them into VGPR
// lanes.)  The arithmetic is that of the generic kernels.
template <typename In, int F, int IO, uint32_t... K>
__device__ __forceinline__ void spec_body(const TraceArgs &A, const In &in) {
  // the host launches these kernels with n * paths < 2^29: indices and
  // every element offset fit 32 bits
  const uint32_t i32 = blockIdx.x * (uint32_t)TRACE_BLOCK + threadIdx.x, n32 = (uint32_t)A.n;
  if (i32 >= n32) return;
  ConstTraceSpec *c_spec = (ConstTraceSpec *)A.spec;
  ConstSurface *c_surf = (ConstSurface *)A.surfaces;
  const TraceSpec &spec = *(const TraceSpec *)c_spec;
  SystemView sys = spec.sys;
  sys.surfaces = A.surfaces;
  const int num_paths = (F & FEAT_RULINGS) ? sys.num_paths : 1;
  Ray<double> r0 = load_io32<IO>(in, i32);
  apply_pre(spec, r0);
  for (int path = 0; path < num_paths; ++path) {
    Ray<double> r = r0;
    int k = 0;
    (propagate_surface<F, K>(*(const optk_surface *)&c_surf[k++], sys, path, r), ...);
    if (A.apply_post) xf_inverse_ray(spec.post, r.px, r.py, r.pz, r.dx, r.dy, r.dz);
    store_io32<IO>(A.out, (uint32_t)path * n32 + i32, r);
  }
}

template <int F, int IO, uint32_t... K>
__global__ __launch_bounds__(TRACE_BLOCK) __attribute__((amdgpu_waves_per_eu(SpecWaves<F>::value)))
void trace_grid_f64_spec(const TraceArgs A, const GridIn g) {
  spec_body<GridIn, F, IO, K...>(A, g);
}

template <int F, int IO, uint32_t... K>
__global__ __launch_bounds__(TRACE_BLOCK) __attribute__((amdgpu_waves_per_eu(SpecWaves<F>::value)))
void trace_rays_f64_spec(const TraceArgs A, const RaysIn in) {
  spec_body<RaysIn, F, IO, K...>(A, in);
}

// the specialised kernels by signature (the order of SIGNATURES): grid
// entry in the sensor-frame I/O layout, ray entry with every output field
typedef void (*SpecGrid64)(TraceArgs, GridIn);
typedef void (*SpecRays64)(TraceArgs, RaysIn);
static const SpecGrid64 SPEC_GRID64[NUM_SIGNATURES] = {
    trace_grid_f64_spec<SIGNATURES[SIG_NEWTONIAN].features, IO_SENSOR, OPTK_SIGNATURE_NEWTONIAN>,
    trace_grid_f64_spec<SIGNATURES[SIG_SPECTROGRAPH].features, IO_SENSOR, OPTK_SIGNATURE_SPECTROGRAPH>,
    trace_grid_f64_spec<SIGNATURES[SIG_SPECTROGRAPH_TILTED].features, IO_SENSOR,
                        OPTK_SIGNATURE_SPECTROGRAPH_TILTED>,
};
static const SpecRays64 SPEC_RAYS64[NUM_SIGNATURES] = {
    trace_rays_f64_spec<SIGNATURES[SIG_NEWTONIAN].features, IO_FULL, OPTK_SIGNATURE_NEWTONIAN>,
    trace_rays_f64_spec<SIGNATURES[SIG_SPECTROGRAPH].features, IO_FULL, OPTK_SIGNATURE_SPECTROGRAPH>,
    trace_rays_f64_spec<SIGNATURES[SIG_SPECTROGRAPH_TILTED].features, IO_FULL,
                        OPTK_SIGNATURE_SPECTROGRAPH_TILTED>,
};

static unsigned spec_blocks(int64_t n) { return (unsigned)((n + TRACE_BLOCK - 1) / TRACE_BLOCK); }

// stop-solve Newton iteration, float64 form ---------------------------------

struct StopParams {
  TraceArgs T;
  optk_surface first;  // sag of the free position's z (FREE_POSITION)
  optk_stop_args a;
};

template <int F>
__global__ __launch_bounds__(256) void stop_newton_f64(const StopParams P) {
  extern __shared__ double staged[];
  TraceSpec *s_spec = reinterpret_cast<TraceSpec *>(staged);
  optk_surface *s_surf = reinterpret_cast<optk_surface *>(s_spec + 1);
  stage_words(s_spec, P.T.spec, sizeof(TraceSpec));
  stage_words(s_surf, P.T.surfaces, sizeof(optk_surface) * (size_t)P.T.num_surfaces);
  __syncthreads();
  SystemView sys = s_spec->sys;
  sys.surfaces = s_surf;
  const optk_stop_args &a = P.a;
  const int64_t n = P.T.n;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    Ray<double> base;
    base.px = a.position[0][i];
    base.py = a.position[1][i];
    base.pz = a.position[2][i];
    base.dx = a.direction[0][i];
    base.dy = a.direction[1][i];
    base.dz = a.direction[2][i];
    base.wavelength = a.wavelength[i];
    base.intensity = 1.0;
    base.attenuation = 0.0;
    base.index = 1.0;
    base.unvignetted = true;
    const double g0 = a.target[0][i], g1 = a.target[1][i];
    double x0 = a.x[0][i], x1 = a.x[1][i];
    int it = 0;
    bool converged = false;
    for (;; ++it) {
      // residual at x, then (unless converged) at x + dx e0 and x + dx e1
      double f[3][2];
      for (int e = 0; e < 3; ++e) {
        const double y0 = e == 1 ? x0 + a.dx : x0;
        const double y1 = e == 2 ? x1 + a.dx : x1;
        Ray<double> r = base;
        if (a.free_component == OPTK_STOP_FREE_DIRECTION) {
          r.dx = y0;
          r.dy = y1;
          r.dz = sqrt(1.0 - (y0 * y0 + y1 * y1));
        } else {
          r.px = y0;
          r.py = y1;
          r.pz = sag_z(P.first, y0, y1, 0.0);
        }
        for (int k = 0; k < P.T.num_surfaces; ++k) propagate_surface<F>(s_surf[k], sys, 0, r);
        if (P.T.apply_post) {
          xf_inverse_ray(s_spec->post, r.px, r.py, r.pz, r.dx, r.dy, r.dz);
        }
        const bool on_position = a.target_component == OPTK_STOP_TARGET_POSITION;
        f[e][0] = (on_position ? r.px : r.dx) - g0;
        f[e][1] = (on_position ? r.py : r.dy) - g1;
        if (e == 0) {
          converged = fabs(f[0][0]) < a.tol && fabs(f[0][1]) < a.tol;
          if (converged || it >= a.max_iter) break;
        }
      }
      if (converged || it >= a.max_iter) break;
      const double j00 = (f[1][0] - f[0][0]) / a.dx, j10 = (f[1][1] - f[0][1]) / a.dx;
      const double j01 = (f[2][0] - f[0][0]) / a.dx, j11 = (f[2][1] - f[0][1]) / a.dx;
      const double det = j00 * j11 - j01 * j10;
      x0 -= (j11 * f[0][0] - j01 * f[0][1]) / det;
      x1 -= (-j10 * f[0][0] + j00 * f[0][1]) / det;
    }
    a.x[0][i] = x0;
    a.x[1][i] = x1;
    if (a.z)
      a.z[i] = a.free_component == OPTK_STOP_FREE_DIRECTION ? sqrt(1.0 - (x0 * x0 + x1 * x1))
                                                           : sag_z(P.first, x0, x1, 0.0);
    a.iterations[i] = converged ? it : a.max_iter + 1;
  }
}


// denormalization of normalized grid coordinates: one thread per value,
// every component in the same pass (reads 8 B per component and the 8 B
// table index, writes 8 B per component) ------------------------------------

__global__ __launch_bounds__(256) void denormalize_kernel(const optk_denormalize_args a, int ncomp,
                                                          int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t w = 0;
  if (a.table_index) {
    w = a.table_index[i];
    w = w < 0 ? 0 : (w >= a.n_table ? a.n_table - 1 : w);
  }
  for (int c = 0; c < ncomp; ++c) {
    const double v = a.in[c][i];
    a.out[c][i] = a.ptp[c][w] * (v + 1.0) / 2.0 + a.lo[c][w];  // _sequential.py:777-785
  }
}


// pupil reductions: one wave per segment, lanes stride the pupil samples,
// then a fixed butterfly over the wave ---------------------------------------

__global__ __launch_bounds__(256) void pupil_reduce_kernel(const optk_pupil_args a, int64_t segments) {
  const int lane = threadIdx.x & 63;
  const int64_t seg = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (seg >= segments) return;  // whole waves leave together
  const int64_t base = seg * a.per_segment;
  double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // count, sx, sy, ax, ay, throughput
  for (int64_t p = lane; p < a.per_segment; p += 64) {
    const int64_t i = base + p;
    const double x = a.x ? a.x[i] : 0.0, y = a.y ? a.y[i] : 0.0;
    v[3] += x;
    v[4] += y;
    if (a.unvignetted[i]) {
      const double w = (a.weight ? a.weight[p] : 1.0) * (a.intensity ? a.intensity[i] : 1.0);
      v[0] += 1.0;
      v[1] += x;
      v[2] += y;
      v[5] += w;
    }
  }
#pragma unroll
  for (int k = 0; k < 6; ++k)
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
  if (lane == 0) {
    if (a.count) a.count[seg] = v[0];
    if (a.sum_x) a.sum_x[seg] = v[1];
    if (a.sum_y) a.sum_y[seg] = v[2];
    if (a.all_x) a.all_x[seg] = v[3];
    if (a.all_y) a.all_y[seg] = v[4];
    if (a.throughput) a.throughput[seg] = v[5];
  }
}


// detector exposure (random numbers and draws in optk_random.h) -----------

__global__ __launch_bounds__(256) void expose_kernel(const optk_expose_args a, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double mean = a.scale * a.photons[i];
    if (!a.noise) {
      a.electrons[i] = mean;
      continue;
    }
    UniformStream rng;
    rng.pixel = (uint64_t)i;
    rng.seed = a.seed;
    double e = poisson_draw(mean, rng);
    if (a.read_noise > 0.0) e += a.read_noise * gaussian_draw(rng);
    a.electrons[i] = e;
  }
}

// batched persistent trace, float64 form ------------------------------------

template <int F> struct Tracer64 {
  const BatchArgs &A;
  const GridIn &in;
  const RaysOut &out;
  TraceSpec *s_spec;
  optk_surface *s_surf;

  __device__ int num_paths() const { return (F & FEAT_RULINGS) ? A.num_paths : 1; }

  __device__ __forceinline__ void stage(int instance) {
    stage_words(s_spec, A.specs64 + instance, sizeof(TraceSpec));
    stage_words(s_surf, A.surfaces + (int64_t)instance * A.num_surfaces,
                sizeof(optk_surface) * (size_t)A.num_surfaces);
  }

  __device__ __forceinline__ void trace(int64_t i, int path, double &x, double &y, double &inten,
                                        bool &unv) {
    Ray<double> r = load_ray(in, i);
    apply_pre(*s_spec, r);
    SystemView sys = s_spec->sys;
    sys.surfaces = s_surf;
    for (int k = 0; k < A.num_surfaces; ++k) propagate_surface<F>(s_surf[k], sys, path, r);
    if (A.apply_post) {
      xf_inverse_ray(s_spec->post, r.px, r.py, r.pz, r.dx, r.dy, r.dz);
    }
    store(out, (int64_t)path * A.n + i, r);
    x = r.px;
    y = r.py;
    inten = r.intensity;
    unv = r.unvignetted;
  }
};

// the same with the instances' surface kinds fixed at compile time (every
// instance of a batch matching one signature): the surfaces unroll as in
// the specialised single-system kernels, the records are read from LDS
template <int F, uint32_t... K> struct Tracer64Spec : Tracer64<F> {
  __device__ __forceinline__ void trace(int64_t i, int path, double &x, double &y, double &inten,
                                        bool &unv) {
    Ray<double> r = load_ray(this->in, i);
    apply_pre(*this->s_spec, r);
    SystemView sys = this->s_spec->sys;
    sys.surfaces = this->s_surf;
    int k = 0;
    (propagate_surface<F, K>(this->s_surf[k++], sys, path, r), ...);
    if (this->A.apply_post) {
      xf_inverse_ray(this->s_spec->post, r.px, r.py, r.pz, r.dx, r.dy, r.dz);
    }
    store(this->out, (int64_t)path * this->A.n + i, r);
    x = r.px;
    y = r.py;
    inten = r.intensity;
    unv = r.unvignetted;
  }
};

template <int F>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(Waves<F>::value)))
void batch_grid_f64(const BatchArgs A, const GridIn in, const RaysOut out) {
  extern __shared__ double staged[];
  Tracer64<F> tr{A, in, out, reinterpret_cast<TraceSpec *>(staged),
                 reinterpret_cast<optk_surface *>(reinterpret_cast<TraceSpec *>(staged) + 1)};
  batch_loop(A, tr);
}

// 2 waves per SIMD (256 VGPRs) for the multilayer layouts, as the generic
// batch kernel: built for 3 or 4 waves (168 / 128 VGPRs, about twice the
// spill slots) this kernel returned every ray vignetted on the tolerance
// batch, while the same source at 256 VGPRs, and with the trace left as a
// call, matches the single-system traces; the narrower builds are not used
template <int F, uint32_t... K>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(Waves<F>::value)))
void batch_grid_f64_spec(const BatchArgs A, const GridIn in, const RaysOut out) {
  extern __shared__ double staged[];
  Tracer64Spec<F, K...> tr{{A, in, out, reinterpret_cast<TraceSpec *>(staged),
                            reinterpret_cast<optk_surface *>(reinterpret_cast<TraceSpec *>(staged) + 1)}};
  batch_loop(A, tr);
}

typedef void (*BatchSpec64)(BatchArgs, GridIn, RaysOut);
// by signature; the geometric Newtonian layout keeps the generic kernel
static const BatchSpec64 BATCH_SPEC64[NUM_SIGNATURES] = {
    nullptr,
    batch_grid_f64_spec<SIGNATURES[SIG_SPECTROGRAPH].features, OPTK_SIGNATURE_SPECTROGRAPH>,
    batch_grid_f64_spec<SIGNATURES[SIG_SPECTROGRAPH_TILTED].features, OPTK_SIGNATURE_SPECTROGRAPH_TILTED>,
};

namespace optk {

size_t batch_lds_f64(int num_surfaces) { return staged_bytes(num_surfaces); }

static void *batch_kernel_f64(int features, int signature) {
  if (signature >= 0 && signature < NUM_SIGNATURES && BATCH_SPEC64[signature])
    return (void *)BATCH_SPEC64[signature];
  if (features == 0) return (void *)batch_grid_f64<0>;
  if (features == (FEAT_TOROID | FEAT_RULINGS)) return (void *)batch_grid_f64<FEAT_TOROID | FEAT_RULINGS>;
  return (void *)batch_grid_f64<FEAT_ALL>;
}

int batch_occupancy_f64(int features, int signature, int block, size_t lds) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_kernel_f64(features, signature), block,
                                                   lds) != hipSuccess)
    return 0;
  return nb;
}

hipError_t batch_launch_f64(int features, int signature, int blocks, int block, const BatchArgs &A,
                            const GridIn &in, const RaysOut &out, hipStream_t stream) {
  size_t lds = staged_bytes(A.num_surfaces);
  if (signature >= 0 && signature < NUM_SIGNATURES && BATCH_SPEC64[signature])
    hipLaunchKernelGGL(BATCH_SPEC64[signature], dim3(blocks), dim3(block), lds, stream, A, in, out);
  else if (features == 0)
    hipLaunchKernelGGL(batch_grid_f64<0>, dim3(blocks), dim3(block), lds, stream, A, in, out);
  else if (features == (FEAT_TOROID | FEAT_RULINGS))
    hipLaunchKernelGGL(batch_grid_f64<FEAT_TOROID | FEAT_RULINGS>, dim3(blocks), dim3(block), lds,
                       stream, A, in, out);
  else
    hipLaunchKernelGGL(batch_grid_f64<FEAT_ALL>, dim3(blocks), dim3(block), lds, stream, A, in, out);
  return hipGetLastError();
}

}  // namespace optk

// per-operator evaluation --------------------------------------------------

struct EvalParams {
  SystemView sys;
  int32_t op, surface, iparam, stack;
  int64_t n;
  const double *in[12];
  double *out[12];
  uint8_t *out_mask;
};

__global__ __launch_bounds__(256) void eval_f64(const EvalParams E) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= E.n) return;
  const optk_surface &s = E.sys.surfaces[E.surface];
  const double *const *in = E.in;
  double *const *out = E.out;
  switch (E.op) {
  case OPTK_OP_SAG_VALUE:
    out[0][i] = sag_z(s, in[0][i], in[1][i], in[2][i]);
    break;
  case OPTK_OP_SAG_NORMAL: {
    V3<double> v = sag_normal(s, in[0][i], in[1][i], in[2][i]);
    out[0][i] = v.x;
    out[1][i] = v.y;
    out[2][i] = v.z;
    break;
  }
  case OPTK_OP_SAG_INTERCEPT: {
    Ray<double> r = {};
    r.px = in[0][i];
    r.py = in[1][i];
    r.pz = in[2][i];
    r.dx = in[3][i];
    r.dy = in[4][i];
    r.dz = in[5][i];
    r.attenuation = in[6][i];
    r.intensity = in[7][i];
    sag_propagate(s, r);
    out[0][i] = r.px;
    out[1][i] = r.py;
    out[2][i] = r.pz;
    out[3][i] = r.dx;
    out[4][i] = r.dy;
    out[5][i] = r.dz;
    out[6][i] = r.intensity;
    break;
  }
  case OPTK_OP_APERTURE:
    E.out_mask[i] = aperture_contains(s, E.sys.sv.tables, in[0][i], in[1][i], in[2][i]) ? 1 : 0;
    break;
  case OPTK_OP_SNELL: {
    V3<double> b = snells_law({in[0][i], in[1][i], in[2][i]}, in[3][i], in[4][i],
                              {in[5][i], in[6][i], in[7][i]}, E.iparam != 0);
    out[0][i] = b.x;
    out[1][i] = b.y;
    out[2][i] = b.z;
    break;
  }
  case OPTK_OP_MULTILAYER: {
    MultilayerOut<double> co;
    MultilayerRT<double> rt = multilayer_efficiency<double>(E.sys.sv, E.stack, in[0][i], C<double>{in[1][i], in[2][i]},
                                            C<double>{in[3][i], in[4][i]}, true, &co);
    out[0][i] = rt.R[0];
    out[1][i] = rt.R[1];
    out[2][i] = rt.T[0];
    out[3][i] = rt.T[1];
    out[4][i] = co.r[0].re;
    out[5][i] = co.r[0].im;
    out[6][i] = co.r[1].re;
    out[7][i] = co.r[1].im;
    out[8][i] = co.t[0].re;
    out[9][i] = co.t[0].im;
    out[10][i] = co.t[1].re;
    out[11][i] = co.t[1].im;
    break;
  }
  case OPTK_OP_SPACING: {
    V3<double> k = spacing_vector<double>(s, {in[0][i], in[1][i], in[2][i]}, {in[3][i], in[4][i], in[5][i]});
    out[0][i] = k.x;
    out[1][i] = k.y;
    out[2][i] = k.z;
    break;
  }
  case OPTK_OP_INCIDENT: {
    V3<double> nrm = {in[5][i], in[6][i], in[7][i]};
    V3<double> kappa = {in[8][i], in[9][i], in[10][i]};
    V3<double> e = incident_effective(s, path_order(s, E.iparam), in[0][i],
                                      {in[1][i], in[2][i], in[3][i]}, in[4][i], nrm, kappa);
    out[0][i] = e.x;
    out[1][i] = e.y;
    out[2][i] = e.z;
    break;
  }
  case OPTK_OP_RULINGS_EFF:
    out[0][i] = rulings_efficiency<double>(s, path_order(s, E.iparam), in[0][i], {in[1][i], in[2][i], in[3][i]},
                                   {in[4][i], in[5][i], in[6][i]}, {in[7][i], in[8][i], in[9][i]},
                                   E.sys.sv.tables);
    break;
  case OPTK_OP_MATERIAL_EFF:
    out[0][i] = material_efficiency(s, E.sys.sv, in[0][i], {in[1][i], in[2][i], in[3][i]}, in[4][i],
                                    in[5][i], {in[6][i], in[7][i], in[8][i]});
    break;
  case OPTK_OP_INDEX:
    out[0][i] = material_index(s, in[0][i], in[1][i]);
    break;
  case OPTK_OP_SNELL_SCALAR: {
    C<double> c = snell_scalar(C<double>{in[0][i], in[1][i]}, C<double>{in[2][i], in[3][i]},
                               C<double>{in[4][i], in[5][i]});
    out[0][i] = c.re;
    out[1][i] = c.im;
    break;
  }
  }
}

// sensor binning --------------------------------------------------------------

struct CollectParams {
  const double *px, *py, *dz, *wavelength, *intensity;
  const float *pxf, *pyf, *dzf, *wavelengthf, *intensityf;
  const uint8_t *unvignetted;
  const double *wedges;
  int32_t nw, nx, ny;
  double x_lo, x_hi, y_lo, y_hi;
  // numpy.linspace's step (hi - lo) / num, and its reciprocal for the first
  // guess of a bin (host-computed once per call)
  double x_step, x_inv, y_step, y_inv;
  double *image, *moment;
  int64_t n;
};

// Bin index of x among numpy.linspace(lo, hi, num + 1) edges with
// numpy.histogramdd's rule (right edge of the last bin inclusive); -1 if out.
// The first guess multiplies by 1/step.  Its error, and the rounding of the
// edges e_j = j * step + lo (e_num = hi) themselves, are a few ulp of
// num * (|lo| + |hi|) / (hi - lo) bins, below one bin while that ratio is
// below 2^48 (the ranges this function serves); two bounded steps each way then
// compare x with the exact edges, so the result is the bin the edges
// define.  The steps compile to selects: no data-dependent loops (which
// had cost the binning kernels their scalar registers).
__device__ __forceinline__ int linspace_bin(double x, double lo, double hi, int num, double step,
                                            double inv) {
  if (!(x >= lo) || !(x <= hi)) return -1;
  if (x == hi) return num - 1;
  int k = (int)floor((x - lo) * inv);
  k = k < 0 ? 0 : (k > num - 1 ? num - 1 : k);
  auto edge = [&](int j) { return j >= num ? hi : (double)j * step + lo; };
#pragma unroll
  for (int r = 0; r < 2; ++r)
    if (k < num - 1 && edge(k + 1) <= x) ++k;
#pragma unroll
  for (int r = 0; r < 2; ++r)
    if (k > 0 && edge(k) > x) --k;
  return k;
}

// The same bin by binary search over the edges formed as numpy forms them
// (e_j = j * step + lo for j < num, e_num = hi; non-decreasing, since each
// of the two roundings is monotone): the largest j < num with e_j <= x.
// Exact for every range, also those far from the origin for their width,
// where the first guess above can miss by more than two bins (the edges may
// even repeat there: zero-width bins, which numpy's rule skips the same way).
__device__ __forceinline__ int linspace_bin_search(double x, double lo, double hi, int num,
                                                   double step) {
  if (!(x >= lo) || !(x <= hi)) return -1;
  if (x == hi) return num - 1;
  int k = 0;
  for (int s = 1 << (31 - __clz(num)); s > 0; s >>= 1)
    if (k + s < num && (double)(k + s) * step + lo <= x) k += s;
  return k;
}

// Bin of x among ascending edges e[0..num] (numpy.searchsorted(e, x,
// 'right') - 1, the last edge inclusive); -1 if out.  Binary search by
// descending powers of two: the trip count depends on num only, so every
// lane of a wave runs the same steps.
__device__ __forceinline__ int edges_bin(double x, const double *e, int num) {
  if (!(x >= e[0]) || !(x <= e[num])) return -1;
  if (x == e[num]) return num - 1;
  int lo = 0;  // largest index in [0, num) with e[lo] <= x
  for (int step = 1 << (31 - __clz(num)); step > 0; step >>= 1)
    if (lo + step < num && e[lo + step] <= x) lo += step;
  return lo;
}

// Tiled binning: a counting sort of the rays into 32 x 32-pixel tiles, then
// one block per tile sums its rays in LDS and adds the tile to the image.
//   count:   block b bins the rays of its chunk once, stores each ray's key
//            (tile << 10 | pixel in tile, -1 when not binned), counts the
//            rays per tile in LDS and writes its row counts[b][*] (coalesced)
//   offsets: one wavefront per tile scans the tile's column of block counts
//            (in place: exclusive offsets) and leaves the tile's total
//   base:    exclusive scan of the totals (one block)
//   scatter: block b reads its chunk's keys and writes one record per ray at
//            base[tile] + offsets[tile][b] + rank: with CT_CHUNK rays per
//            block a tile receives runs of ~CT_CHUNK / tiles records, whole
//            cache lines for a 2048^2 image
//   tiles:   blocks sum a tile's rays with LDS float64 atomics and add its
//            pixels to the image (one block per tile, or CT_SPLIT blocks and
//            float64 atomics for a tile holding many rays)
// Records: float32 rays pack (pixel, weight, cosine) into one 16-byte
// store; float64 rays keep a 2-byte pixel array beside (weight, cosine).
// The moment weight x cosine is formed in float64 in the tiles kernel, as
// the one-pass kernel forms it.
constexpr int CT_SIDE = 32;
constexpr int CT_PIX = CT_SIDE * CT_SIDE;
constexpr int CT_MAX_TILES = 16384;
constexpr int64_t CT_CHUNK = 32768;  // rays per block in count / scatter
constexpr int CT_THREADS = 1024;     // threads per count / scatter block

struct RecordF32 {
  uint32_t pix;
  float inten, dz;
  uint32_t pad;
};

struct TiledParams {
  CollectParams C;
  int32_t tiles_x, tiles_y, num_tiles, blocks;
  int32_t *counts;      // [blocks][num_tiles], then exclusive offsets in place
  int32_t *base;        // [num_tiles + 1]
  int32_t *keys;        // [n] tile << 10 | pixel, -1 when not binned
  uint16_t *sorted_pix; // float64 rays
  double2 *sorted_ab;   // float64 rays: (weight, cosine)
  RecordF32 *sorted32;  // float32 rays
};

constexpr int CT_UNROLL = 4;  // rays per thread per step of scatter / tiles

// position, wavelength and mask of one ray, for the key
struct BinIn {
  double x, y, w;
  bool live;
};

template <bool F32> __device__ __forceinline__ BinIn bin_load(const CollectParams &C, int64_t i) {
  BinIn b;
  b.live = !(C.unvignetted && !C.unvignetted[i]);
  if (F32) {
    b.x = C.pxf[i];
    b.y = C.pyf[i];
    b.w = C.wavelengthf[i];
  } else {
    b.x = C.px[i];
    b.y = C.py[i];
    b.w = C.wavelength[i];
  }
  return b;
}

// tile << 10 | pixel in tile, or -1 when the ray is masked or outside
__device__ __forceinline__ int32_t bin_key(const TiledParams &T, const BinIn &b) {
  if (!b.live) return -1;
  const int bw = edges_bin(b.w, T.C.wedges, T.C.nw);
  const int bx = linspace_bin(b.x, T.C.x_lo, T.C.x_hi, T.C.nx, T.C.x_step, T.C.x_inv);
  const int by = linspace_bin(b.y, T.C.y_lo, T.C.y_hi, T.C.ny, T.C.y_step, T.C.y_inv);
  if (bw < 0 || bx < 0 || by < 0) return -1;
  const int tile = (bw * T.tiles_x + bx / CT_SIDE) * T.tiles_y + by / CT_SIDE;
  return (tile << 10) | ((bx % CT_SIDE) * CT_SIDE + by % CT_SIDE);
}

// count and scatter hold one 32-bit word per tile in dynamic LDS; both are
// specialised by the rays' precision (one set of array pointers live)
template <bool F32>
__global__ __launch_bounds__(CT_THREADS) void collect_count(const TiledParams T) {
  extern __shared__ int32_t cnt[];
  for (int t = threadIdx.x; t < T.num_tiles; t += CT_THREADS) cnt[t] = 0;
  __syncthreads();
  const int64_t lo = (int64_t)blockIdx.x * CT_CHUNK;
  const int64_t hi = lo + CT_CHUNK < T.C.n ? lo + CT_CHUNK : T.C.n;
  // one ray per thread per step (unrolling the binning's rare exact walks
  // would exhaust the scalar registers that hold their exec masks)
  for (int64_t i = lo + threadIdx.x; i < hi; i += CT_THREADS) {
    const int32_t key = bin_key(T, bin_load<F32>(T.C, i));
    if (key >= 0) atomicAdd(&cnt[key >> 10], 1);
    T.keys[i] = key;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < T.num_tiles; t += CT_THREADS)
    T.counts[(int64_t)blockIdx.x * T.num_tiles + t] = cnt[t];
}

__global__ __launch_bounds__(256) void collect_offsets(const TiledParams T) {
  const int lane = threadIdx.x & 63;
  const int t = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (t >= T.num_tiles) return;  // whole wavefronts
  int32_t *col = T.counts + t;  // entry b at col[b * num_tiles]
  int32_t carry = 0;
  for (int b0 = 0; b0 < T.blocks; b0 += 64) {
    const int b = b0 + lane;
    const int64_t k = (int64_t)b * T.num_tiles;
    const int32_t c = b < T.blocks ? col[k] : 0;
    int32_t incl = c;
    for (int off = 1; off < 64; off <<= 1) {
      const int32_t v = __shfl_up(incl, off, 64);
      if (lane >= off) incl += v;
    }
    if (b < T.blocks) col[k] = carry + incl - c;
    carry += __shfl(incl, 63, 64);
  }
  if (lane == 0) T.base[t] = carry;  // the tile's total, scanned by collect_base
}

__global__ __launch_bounds__(1024) void collect_base(const TiledParams T) {
  // exclusive scan of T.base[0 .. num_tiles) in place (num_tiles <= 16384):
  // each thread scans 16 entries, then a block scan of the thread sums
  constexpr int PER = CT_MAX_TILES / 1024;
  __shared__ int32_t part[1024];
  const int t = threadIdx.x;
  int32_t v[PER], sum = 0;
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int k = PER * t + j;
    v[j] = k < T.num_tiles ? T.base[k] : 0;
    sum += v[j];
  }
  part[t] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int32_t add = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += add;
    __syncthreads();
  }
  int32_t run = part[t] - sum;  // exclusive prefix of this thread's entries
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int k = PER * t + j;
    if (k < T.num_tiles) T.base[k] = run;
    run += v[j];
  }
  if (t == 1023) T.base[T.num_tiles] = part[1023];
}

template <bool F32>
__global__ __launch_bounds__(CT_THREADS) void collect_scatter(const TiledParams T) {
  extern __shared__ int32_t cursor[];
  for (int t = threadIdx.x; t < T.num_tiles; t += CT_THREADS)
    cursor[t] = T.base[t] + T.counts[(int64_t)blockIdx.x * T.num_tiles + t];
  __syncthreads();
  const int64_t lo = (int64_t)blockIdx.x * CT_CHUNK;
  const int64_t hi = lo + CT_CHUNK < T.C.n ? lo + CT_CHUNK : T.C.n;
  for (int64_t i0 = lo + threadIdx.x; i0 < hi; i0 += CT_UNROLL * CT_THREADS) {
    int32_t key[CT_UNROLL];
    double a[CT_UNROLL], c[CT_UNROLL];
#pragma unroll
    for (int u = 0; u < CT_UNROLL; ++u) {
      const int64_t i = i0 + (int64_t)u * CT_THREADS;
      key[u] = -1;
      a[u] = c[u] = 0.0;
      if (i < hi) {
        key[u] = T.keys[i];
        a[u] = F32 ? (double)T.C.intensityf[i] : T.C.intensity[i];
        c[u] = F32 ? (double)T.C.dzf[i] : T.C.dz[i];
      }
    }
#pragma unroll
    for (int u = 0; u < CT_UNROLL; ++u) {
      if (key[u] < 0) continue;
      const int32_t slot = atomicAdd(&cursor[key[u] >> 10], 1);
      if (F32) {
        T.sorted32[slot] = RecordF32{(uint32_t)(key[u] & 1023), (float)a[u], (float)c[u], 0u};
      } else {
        T.sorted_pix[slot] = (uint16_t)(key[u] & 1023);
        T.sorted_ab[slot] = double2{a[u], c[u]};
      }
    }
  }
}

// Images of at most CT_STAGED_TILES tiles: the scatter stages each
// sub-chunk of CT_SUB rays in LDS in tile order (a local counting sort) and
// then writes it out, so that one tile's records leave the block as
// consecutive stores, whole cache lines, instead of one scattered store per
// ray that concurrent blocks complete piecewise.  Float32 rays stage their
// 16-byte records; float64 rays stage (weight, cosine) and the 2-byte pixel.
//   1. zero the local counts            2. load, rank within the tile (LDS atomic)
//   3. scan the counts (local offsets)  4. place records and global slots in LDS
//   5. write out in local order, advance the tiles' cursors
constexpr int CT_SUB = CT_UNROLL * CT_THREADS;  // rays per sub-chunk
constexpr int CT_STAGED_TILES = 4096;

template <bool F32> struct StageRecord;
template <> struct StageRecord<true> {
  RecordF32 r;
};
template <> struct StageRecord<false> {
  double2 ab;
};

template <bool F32>
__global__ __launch_bounds__(CT_THREADS) void collect_scatter_staged(const TiledParams T) {
  extern __shared__ int32_t dyn[];
  int32_t *cursor = dyn;                    // [num_tiles] next global slot of each tile
  int32_t *loff = dyn + T.num_tiles;        // [num_tiles] local counts, then local offsets
  int32_t *lcnt = dyn + 2 * T.num_tiles;    // [num_tiles] local counts of the pass
  __shared__ StageRecord<F32> stage[CT_SUB];
  __shared__ uint16_t stage_pix[F32 ? 1 : CT_SUB];
  __shared__ int32_t dest[CT_SUB];
  __shared__ int32_t wsum[CT_THREADS / 64];
  __shared__ int32_t total;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int per = (T.num_tiles + CT_THREADS - 1) / CT_THREADS;  // tiles scanned per thread
  const int t0 = tid * per;
  for (int t = tid; t < T.num_tiles; t += CT_THREADS)
    cursor[t] = T.base[t] + T.counts[(int64_t)blockIdx.x * T.num_tiles + t];
  const int64_t lo = (int64_t)blockIdx.x * CT_CHUNK;
  const int64_t hi = lo + CT_CHUNK < T.C.n ? lo + CT_CHUNK : T.C.n;
  for (int64_t s0 = lo; s0 < hi; s0 += CT_SUB) {
    for (int t = tid; t < T.num_tiles; t += CT_THREADS) loff[t] = 0;
    __syncthreads();
    typedef typename std::conditional<F32, float, double>::type Real;
    int32_t key[CT_UNROLL], rank[CT_UNROLL];
    Real a[CT_UNROLL], c[CT_UNROLL];
#pragma unroll
    for (int u = 0; u < CT_UNROLL; ++u) {
      const int64_t i = s0 + (int64_t)u * CT_THREADS + tid;
      key[u] = -1;
      a[u] = c[u] = Real(0);
      if (i < hi) {
        key[u] = T.keys[i];
        if constexpr (F32) {
          a[u] = T.C.intensityf[i];
          c[u] = T.C.dzf[i];
        } else {
          a[u] = T.C.intensity[i];
          c[u] = T.C.dz[i];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < CT_UNROLL; ++u) rank[u] = key[u] >= 0 ? atomicAdd(&loff[key[u] >> 10], 1) : 0;
    __syncthreads();
    // exclusive scan of loff: per-thread runs of `per` tiles, wave scan, scan of the wave sums
    int32_t sum = 0;
    for (int j = 0; j < per; ++j) {
      const int t = t0 + j;
      if (t < T.num_tiles) sum += loff[t];
    }
    int32_t incl = sum;
    for (int off = 1; off < 64; off <<= 1) {
      const int32_t v = __shfl_up(incl, off, 64);
      if (lane >= off) incl += v;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    if (wave == 0) {
      const int32_t w = lane < CT_THREADS / 64 ? wsum[lane] : 0;
      int32_t wi = w;
      for (int off = 1; off < 64; off <<= 1) {
        const int32_t v = __shfl_up(wi, off, 64);
        if (lane >= off) wi += v;
      }
      if (lane < CT_THREADS / 64) wsum[lane] = wi - w;
      if (lane == 63) total = wi;
    }
    __syncthreads();
    int32_t run = wsum[wave] + incl - sum;
    for (int j = 0; j < per; ++j) {
      const int t = t0 + j;
      if (t < T.num_tiles) {
        const int32_t cnt = loff[t];
        lcnt[t] = cnt;
        loff[t] = run;
        run += cnt;
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < CT_UNROLL; ++u) {
      if (key[u] < 0) continue;
      const int tile = key[u] >> 10;
      const int32_t p = loff[tile] + rank[u];
      if constexpr (F32) {
        stage[p].r = RecordF32{(uint32_t)(key[u] & 1023), a[u], c[u], 0u};
      } else {
        stage[p].ab = double2{a[u], c[u]};
        stage_pix[p] = (uint16_t)(key[u] & 1023);
      }
      dest[p] = cursor[tile] + rank[u];
    }
    __syncthreads();
    const int32_t n_sub = total;
    for (int32_t p = tid; p < n_sub; p += CT_THREADS) {
      if constexpr (F32) {
        T.sorted32[dest[p]] = stage[p].r;
      } else {
        T.sorted_ab[dest[p]] = stage[p].ab;
        T.sorted_pix[dest[p]] = stage_pix[p];
      }
    }
    for (int j = 0; j < per; ++j) {
      const int t = t0 + j;
      if (t < T.num_tiles) cursor[t] += lcnt[t];
    }
  }
}

// Tiles kernel: CT_SPLIT blocks per tile (blockIdx.x = tile * CT_SPLIT +
// part).  A tile with at most CT_SPLIT_MIN rays is summed by part 0 alone,
// which adds its pixels to the image with plain read-modify-writes; a
// fuller tile (images concentrate the rays in a few tiles) is split into
// CT_SPLIT equal record ranges whose partial sums are added with float64
// atomics.  Every pixel of a tile is thus written one way only.
constexpr int CT_SPLIT = 4;
constexpr int32_t CT_SPLIT_MIN = 4096;

template <bool F32>
__global__ __launch_bounds__(256) void collect_tiles(const TiledParams T) {
  // launched with 256 threads
  __shared__ double acc_a[CT_PIX];
  __shared__ double acc_b[CT_PIX];
  const int tile = blockIdx.x / CT_SPLIT, part = blockIdx.x % CT_SPLIT;
  const int32_t t_lo = T.base[tile], t_hi = T.base[tile + 1];
  const bool split = t_hi - t_lo > CT_SPLIT_MIN;
  if (!split && part != 0) return;  // whole block
  const int32_t len = t_hi - t_lo;
  const int32_t lo = split ? t_lo + (int32_t)((int64_t)len * part / CT_SPLIT) : t_lo;
  const int32_t hi = split ? t_lo + (int32_t)((int64_t)len * (part + 1) / CT_SPLIT) : t_hi;
  if (hi == lo) return;  // whole block: nothing to add
  for (int p = threadIdx.x; p < CT_PIX; p += blockDim.x) {
    acc_a[p] = 0.0;
    acc_b[p] = 0.0;
  }
  __syncthreads();
  for (int32_t j0 = lo + threadIdx.x; j0 < hi; j0 += CT_UNROLL * 256) {
    int pix[CT_UNROLL];
    double a[CT_UNROLL], c[CT_UNROLL];
#pragma unroll
    for (int u = 0; u < CT_UNROLL; ++u) {
      const int32_t j = j0 + u * 256;
      pix[u] = -1;
      if (j < hi) {
        if (F32) {
          const RecordF32 r = T.sorted32[j];
          pix[u] = (int)r.pix;
          a[u] = r.inten;
          c[u] = r.dz;
        } else {
          pix[u] = T.sorted_pix[j];
          const double2 ab = T.sorted_ab[j];
          a[u] = ab.x;
          c[u] = ab.y;
        }
      }
    }
#pragma unroll
    for (int u = 0; u < CT_UNROLL; ++u) {
      if (pix[u] < 0) continue;
      atomicAdd(&acc_a[pix[u]], a[u]);
      atomicAdd(&acc_b[pix[u]], a[u] * c[u]);
    }
  }
  __syncthreads();
  const int w = tile / (T.tiles_x * T.tiles_y);
  const int r = tile - w * T.tiles_x * T.tiles_y;
  const int tx = r / T.tiles_y, ty = r - tx * T.tiles_y;
  for (int p = threadIdx.x; p < CT_PIX; p += blockDim.x) {
    const int gx = tx * CT_SIDE + p / CT_SIDE, gy = ty * CT_SIDE + p % CT_SIDE;
    if (gx >= T.C.nx || gy >= T.C.ny) continue;
    const int64_t k = ((int64_t)w * T.C.nx + gx) * T.C.ny + gy;
    if (split) {
      if (acc_a[p] != 0.0 || acc_b[p] != 0.0) {
        unsafeAtomicAdd(&T.C.image[k], acc_a[p]);
        unsafeAtomicAdd(&T.C.moment[k], acc_b[p]);
      }
    } else {
      T.C.image[k] += acc_a[p];
      T.C.moment[k] += acc_b[p];
    }
  }
}

// One ray per thread.  Rays of a wavefront that land in the same bin are
// combined before the atomic: the wave picks the first pending lane's bin,
// sums the weights of every lane in that bin with shuffles, and one lane
// adds the sums.  Images of point sources (many rays per pixel in a row)
// then cost one atomic per pixel per wave instead of one per ray; when a
// round finds fewer than 4 lanes in a bin the remaining lanes (scattered
// rays) use their own atomics.
template <bool SEARCH>
__global__ __launch_bounds__(256) void collect_kernel(const CollectParams C) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  int64_t idx = -1;
  double inten = 0.0, dz = 0.0;
  if (i < C.n && !(C.unvignetted && !C.unvignetted[i])) {
    double x, y, w;
    if (C.px) {
      x = C.px[i];
      y = C.py[i];
      dz = C.dz[i];
      w = C.wavelength[i];
      inten = C.intensity[i];
    } else {
      x = C.pxf[i];
      y = C.pyf[i];
      dz = C.dzf[i];
      w = C.wavelengthf[i];
      inten = C.intensityf[i];
    }
    int bw = edges_bin(w, C.wedges, C.nw);
    int bx = SEARCH ? linspace_bin_search(x, C.x_lo, C.x_hi, C.nx, C.x_step)
                    : linspace_bin(x, C.x_lo, C.x_hi, C.nx, C.x_step, C.x_inv);
    int by = SEARCH ? linspace_bin_search(y, C.y_lo, C.y_hi, C.ny, C.y_step)
                    : linspace_bin(y, C.y_lo, C.y_hi, C.ny, C.y_step, C.y_inv);
    if (bw >= 0 && bx >= 0 && by >= 0) idx = ((int64_t)bw * C.nx + bx) * C.ny + by;
  }
  uint64_t pending = __ballot(idx >= 0);
  while (pending) {
    const int leader = __ffsll((long long)pending) - 1;
    const int64_t key = __shfl(idx, leader, 64);
    const bool mine = idx == key && ((pending >> lane) & 1ull);
    const uint64_t group = __ballot(mine);
    if (__popcll(group) < 4) break;  // scattered: per-lane atomics below
    double a = mine ? inten : 0.0;
    double b = mine ? inten * dz : 0.0;
    for (int off = 32; off > 0; off >>= 1) {
      a += __shfl_xor(a, off, 64);
      b += __shfl_xor(b, off, 64);
    }
    if (lane == leader) {
      unsafeAtomicAdd(&C.image[key], a);
      unsafeAtomicAdd(&C.moment[key], b);
    }
    pending &= ~group;
  }
  if ((pending >> lane) & 1ull) {
    unsafeAtomicAdd(&C.image[idx], inten);
    unsafeAtomicAdd(&C.moment[idx], inten * dz);
  }
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

static SystemView view_of(const optk_system *s) {
  SystemView v;
  v.surfaces = s->surfaces;
  v.num_surfaces = s->num_surfaces;
  v.num_paths = s->num_paths;
  v.sv.grids = s->grids;
  v.sv.grid_lo = s->grid_lo;
  v.sv.layers = s->layers;
  v.sv.stacks = s->stacks;
  v.sv.chemicals = s->chemicals;
  v.sv.tables = s->tables;
  return v;
}

template <typename T> static int32_t upload(T **dst, const T *src, int64_t count) {
  *dst = nullptr;
  if (count <= 0) return OPTK_OK;
  if (!src) return fail(OPTK_EINVAL, "null host array for a non-empty record list");
  hipError_t e = hipMalloc((void **)dst, sizeof(T) * count);
  if (e != hipSuccess) return fail(OPTK_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  e = hipMemcpy(*dst, src, sizeof(T) * count, hipMemcpyHostToDevice);
  if (e != hipSuccess) return fail(OPTK_EDEVICE, std::string("hipMemcpy: ") + hipGetErrorString(e));
  return OPTK_OK;
}

static bool valid_affine(const optk_affine &a) {
  return a.kind >= OPTK_XF_IDENTITY && a.kind <= OPTK_XF_AFFINE;
}

// Check every index a kernel will follow, so no kernel can read outside the
// uploaded records.
namespace optk {
int32_t validate_desc(const optk_system_desc *d) {
  // every surface record is staged in LDS (about 0.9 KiB each)
  if (d->num_surfaces < 0 || d->num_surfaces > OPTK_MAX_SURFACES)
    return fail(OPTK_EINVAL, "num_surfaces out of range (at most 64)");
  if (d->num_paths < 1 || d->num_paths > OPTK_MAX_ORDERS) return fail(OPTK_EINVAL, "num_paths out of range");
  for (int c = 0; c < d->num_chemicals; ++c) {
    const optk_chemical &ch = d->chemicals[c];
    if (ch.count < 2 || ch.offset < 0 || (int64_t)ch.offset + 3LL * ch.count > d->num_tables)
      return fail(OPTK_EINVAL, "chemical table out of range");
  }
  for (int l = 0; l < d->num_layers; ++l) {
    const optk_layer &L = d->layers[l];
    if (L.chemical < -1 || L.chemical >= d->num_chemicals) return fail(OPTK_EINVAL, "layer chemical index out of range");
  }
  for (int k = 0; k < d->num_stacks; ++k) {
    const optk_stack &st = d->stacks[k];
    if (st.num_blocks < 0 || st.num_blocks > OPTK_MAX_BLOCKS) return fail(OPTK_EINVAL, "stack block count out of range");
    if (st.substrate_layer < -1 || st.substrate_layer >= d->num_layers) return fail(OPTK_EINVAL, "substrate index out of range");
    for (int b = 0; b < st.num_blocks; ++b) {
      const optk_block &bl = st.blocks[b];
      if (bl.first_layer < 0 || bl.num_layers < 0 || bl.first_layer + bl.num_layers > d->num_layers)
        return fail(OPTK_EINVAL, "stack block layers out of range");
      if (bl.kind == OPTK_BLOCK_PERIODIC && bl.num_periods < 1) return fail(OPTK_EINVAL, "num_periods must be >= 1");
    }
  }
  for (int i = 0; i < d->num_surfaces; ++i) {
    const optk_surface &s = d->surfaces[i];
    if (!valid_affine(s.transform) || !valid_affine(s.sag_transform) || !valid_affine(s.spacing_transform) ||
        !valid_affine(s.aperture_transform))
      return fail(OPTK_EINVAL, "bad transformation kind");
    if (s.sag_kind < OPTK_SAG_NONE || s.sag_kind > OPTK_SAG_TOROID) return fail(OPTK_EINVAL, "bad sag kind");
    if (s.material_kind < OPTK_MATERIAL_VACUUM || s.material_kind > OPTK_MATERIAL_SILICON_SENSOR)
      return fail(OPTK_EINVAL, "bad material kind");
    if (s.material_kind == OPTK_MATERIAL_MEASURED_MIRROR &&
        (s.material_table < 0 || s.material_table_len < 2 ||
         (int64_t)s.material_table + 2LL * s.material_table_len > d->num_tables))
      return fail(OPTK_EINVAL, "measured-mirror table out of range");
    if ((s.material_kind == OPTK_MATERIAL_MULTILAYER_MIRROR || s.material_kind == OPTK_MATERIAL_MULTILAYER_FILM ||
         s.material_kind == OPTK_MATERIAL_SILICON_SENSOR) &&
        (s.stack < 0 || s.stack >= d->num_stacks))
      return fail(OPTK_EINVAL, "multilayer stack index out of range");
    if (s.rulings_kind < OPTK_RULINGS_NONE || s.rulings_kind > OPTK_RULINGS_RECTANGULAR)
      return fail(OPTK_EINVAL, "bad rulings kind");
    if (s.rulings_kind != OPTK_RULINGS_NONE) {
      if (s.num_orders != 1 && s.num_orders != d->num_paths) return fail(OPTK_EINVAL, "diffraction-order count mismatch");
      if (s.spacing_kind < 0 || s.spacing_kind > OPTK_SPACING_HOLOGRAPHIC) return fail(OPTK_EINVAL, "bad spacing kind");
      if (s.num_terms < 0 || s.num_terms > OPTK_MAX_POLY) return fail(OPTK_EINVAL, "bad polynomial term count");
      if (s.rulings_kind == OPTK_RULINGS_MEASURED &&
          (s.rulings_table < 0 || s.rulings_table_len < 2 ||
           (int64_t)s.rulings_table + 2LL * s.rulings_table_len > d->num_tables))
        return fail(OPTK_EINVAL, "measured-rulings table out of range");
    }
    if (s.aperture_kind < OPTK_APERTURE_NONE || s.aperture_kind > OPTK_APERTURE_POLYGON)
      return fail(OPTK_EINVAL, "bad aperture kind");
    if (s.aperture_kind == OPTK_APERTURE_POLYGON &&
        (s.polygon_count < 3 || s.polygon_offset < 0 ||
         (int64_t)s.polygon_offset + 2LL * s.polygon_count > d->num_tables))
      return fail(OPTK_EINVAL, "polygon table out of range");
  }
  for (int k = 0; k < 2; ++k)
    if (!valid_affine(d->pre[k])) return fail(OPTK_EINVAL, "bad pre transformation");
  if (!valid_affine(d->post)) return fail(OPTK_EINVAL, "bad post transformation");
  return OPTK_OK;
}

int32_t system_features(const optk_system_desc *d) { return variant_of(features_of(d)); }

int32_t upload_chem_grids(const optk_chemical *chems, int num_chemicals, const double *tables,
                          ChemGrid **grids, int32_t **grid_lo) {
  *grids = nullptr;
  *grid_lo = nullptr;
  if (num_chemicals <= 0) return OPTK_OK;
  int64_t count = 0;
  build_chem_grids(chems, num_chemicals, tables, nullptr, nullptr, &count);
  std::vector<ChemGrid> g(num_chemicals);
  if (count == 0) count = 1;  // no chemical has a grid: one unused entry
  std::vector<int32_t> lo(count, 0);
  build_chem_grids(chems, num_chemicals, tables, g.data(), lo.data(), nullptr);
  hipError_t e = hipMalloc((void **)grids, sizeof(ChemGrid) * num_chemicals);
  if (e == hipSuccess) e = hipMalloc((void **)grid_lo, sizeof(int32_t) * count);
  if (e != hipSuccess) return fail(OPTK_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  e = hipMemcpy(*grids, g.data(), sizeof(ChemGrid) * num_chemicals, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(*grid_lo, lo.data(), sizeof(int32_t) * count, hipMemcpyHostToDevice);
  if (e != hipSuccess) return fail(OPTK_EDEVICE, std::string("hipMemcpy: ") + hipGetErrorString(e));
  return OPTK_OK;
}
}  // namespace optk

static_assert(sizeof(optk_surface) % 8 == 0 && sizeof(TraceSpec) % 8 == 0, "records are staged as words");

static dim3 grid_for(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

extern "C" {

int32_t optk_abi_version(void) { return OPTK_ABI_VERSION; }

const char *optk_last_error(void) { return g_last_error.c_str(); }

int64_t optk_sizeof(const char *name) {
  if (!name) return -1;
  std::string s(name);
  if (s == "optk_affine") return sizeof(optk_affine);
  if (s == "optk_surface") return sizeof(optk_surface);
  if (s == "optk_layer") return sizeof(optk_layer);
  if (s == "optk_block") return sizeof(optk_block);
  if (s == "optk_stack") return sizeof(optk_stack);
  if (s == "optk_chemical") return sizeof(optk_chemical);
  if (s == "optk_system_desc") return sizeof(optk_system_desc);
  if (s == "optk_rays") return sizeof(optk_rays);
  if (s == "optk_grid") return sizeof(optk_grid);
  if (s == "optk_eval_args") return sizeof(optk_eval_args);
  if (s == "optk_collect_args") return sizeof(optk_collect_args);
  if (s == "optk_batch_desc") return sizeof(optk_batch_desc);
  if (s == "optk_tune") return sizeof(optk_tune);
  if (s == "optk_mlp_args") return sizeof(optk_mlp_args);
  if (s == "optk_surrogate_desc") return sizeof(optk_surrogate_desc);
  if (s == "optk_surrogate_io") return sizeof(optk_surrogate_io);
  if (s == "optk_stop_args") return sizeof(optk_stop_args);
  if (s == "optk_pupil_args") return sizeof(optk_pupil_args);
  if (s == "optk_denormalize_args") return sizeof(optk_denormalize_args);
  if (s == "optk_expose_args") return sizeof(optk_expose_args);
  return -1;
}

int32_t optk_system_destroy(optk_system *sys) {
  if (!sys) return OPTK_OK;
  (void)hipFree(sys->surfaces);
  (void)hipFree(sys->layers);
  (void)hipFree(sys->stacks);
  (void)hipFree(sys->chemicals);
  (void)hipFree(sys->tables);
  (void)hipFree(sys->spec);
  (void)hipFree(sys->fast);
  (void)hipFree(sys->fast_spec);
  (void)hipFree(sys->tables32);
  (void)hipFree(sys->grids);
  (void)hipFree(sys->grid_lo);
  delete sys;
  return OPTK_OK;
}

int32_t optk_system_create(const optk_system_desc *desc, optk_system **out) {
  if (!desc || !out) return fail(OPTK_EINVAL, "null argument");
  *out = nullptr;
  if ((desc->num_surfaces > 0 && !desc->surfaces) || (desc->num_layers > 0 && !desc->layers) ||
      (desc->num_stacks > 0 && !desc->stacks) || (desc->num_chemicals > 0 && !desc->chemicals) ||
      (desc->num_tables > 0 && !desc->tables))
    return fail(OPTK_EINVAL, "null record array");
  int32_t rc = validate_desc(desc);
  if (rc) return rc;
  optk_system *s = new optk_system();
  s->num_surfaces = desc->num_surfaces;
  s->num_paths = desc->num_paths;
  s->num_layers = desc->num_layers;
  s->num_stacks = desc->num_stacks;
  s->num_chemicals = desc->num_chemicals;
  s->num_tables = desc->num_tables;
  s->pre[0] = desc->pre[0];
  s->pre[1] = desc->pre[1];
  s->pre_is_inverse[0] = desc->pre_is_inverse[0];
  s->pre_is_inverse[1] = desc->pre_is_inverse[1];
  s->post = desc->post;
  s->host_surfaces.assign(desc->surfaces, desc->surfaces + desc->num_surfaces);
  for (auto &surf : s->host_surfaces) derive_constants(surf);
  s->features = variant_of(features_of(desc));
  s->signature = find_signature(s->host_surfaces.data(), desc->num_surfaces, s->features);
  if ((rc = upload(&s->surfaces, s->host_surfaces.data(), desc->num_surfaces)) ||
      (rc = upload(&s->layers, desc->layers, desc->num_layers)) ||
      (rc = upload(&s->stacks, desc->stacks, desc->num_stacks)) ||
      (rc = upload(&s->chemicals, desc->chemicals, desc->num_chemicals)) ||
      (rc = upload(&s->tables, desc->tables, desc->num_tables)) ||
      (rc = upload_chem_grids(desc->chemicals, desc->num_chemicals, desc->tables, &s->grids,
                              &s->grid_lo))) {
    optk_system_destroy(s);
    return rc;
  }
  TraceSpec spec;
  std::memset(&spec, 0, sizeof(spec));
  spec.sys = view_of(s);
  spec.pre[0] = s->pre[0];
  spec.pre[1] = s->pre[1];
  spec.pre_is_inverse[0] = s->pre_is_inverse[0];
  spec.pre_is_inverse[1] = s->pre_is_inverse[1];
  spec.post = s->post;
  spec.apply_post = 0;
  if ((rc = upload(&s->spec, &spec, 1))) {
    optk_system_destroy(s);
    return rc;
  }
  // float32 form: composed transformations and float32 tables
  std::vector<FastSurface> fast(desc->num_surfaces);
  FastSpec fast_spec;
  std::memset(&fast_spec, 0, sizeof(fast_spec));
  fast_build(s->host_surfaces.data(), desc->num_surfaces, desc->pre, desc->pre_is_inverse,
             desc->post, fast.data(), fast_spec);
  std::vector<float> tables32(desc->tables, desc->tables + desc->num_tables);
  if ((rc = upload(&s->fast, fast.data(), desc->num_surfaces)) ||
      (rc = upload(&s->fast_spec, &fast_spec, 1)) ||
      (rc = upload(&s->tables32, tables32.data(), desc->num_tables))) {
    optk_system_destroy(s);
    return rc;
  }
  *out = s;
  return OPTK_OK;
}

int32_t optk_system_specialised(const optk_system *sys) { return sys ? sys->signature : -1; }

static void fill_out(RaysOut &o, const optk_rays *r) {
  std::memset(&o, 0, sizeof(o));
  if (!r) return;
  for (int k = 0; k < 3; ++k) {
    o.p[k] = (double *)r->position[k];
    o.d[k] = (double *)r->direction[k];
  }
  o.wavelength = (double *)r->wavelength;
  o.intensity = (double *)r->intensity;
  o.attenuation = (double *)r->attenuation;
  o.index = (double *)r->index_refraction;
  o.unvignetted = r->unvignetted;
}

static void fill_out32(FRaysOut &o, const optk_rays *r) {
  std::memset(&o, 0, sizeof(o));
  if (!r) return;
  for (int k = 0; k < 3; ++k) {
    o.p[k] = (float *)r->position[k];
    o.d[k] = (float *)r->direction[k];
  }
  o.wavelength = (float *)r->wavelength;
  o.intensity = (float *)r->intensity;
  o.attenuation = (float *)r->attenuation;
  o.index = (float *)r->index_refraction;
  o.unvignetted = r->unvignetted;
}

static int32_t prepare(const optk_system *sys, int32_t precision, const optk_rays *out,
                       const optk_rays *accum, int64_t n, int32_t flags, TraceArgs &A,
                       RaysOut &acc) {
  if (!sys) return fail(OPTK_EINVAL, "null system");
  if (!out) return fail(OPTK_EINVAL, "null output rays");
  if (n < 0) return fail(OPTK_EINVAL, "negative ray count");
  if (precision != OPTK_F64 && precision != OPTK_F32) return fail(OPTK_EINVAL, "bad precision");
  std::memset(&A, 0, sizeof(A));
  A.spec = sys->spec;
  A.surfaces = sys->surfaces;
  A.num_surfaces = sys->num_surfaces;
  A.n = n;
  A.apply_post = (flags & OPTK_TRACE_APPLY_POST) ? 1 : 0;
  fill_out(A.out, out);
  fill_out(acc, accum);
  return OPTK_OK;
}

int32_t optk_trace(const optk_system *sys, int32_t precision, const optk_rays *in,
                   const optk_rays *out, const optk_rays *accum, int64_t n, int32_t flags,
                   void *hip_stream) {
  TraceArgs A;
  RaysOut acc;
  int32_t rc = prepare(sys, precision, out, accum, n, flags, A, acc);
  if (rc) return rc;
  if (!in) return fail(OPTK_EINVAL, "null input rays");
  for (int k = 0; k < 3; ++k)
    if (!in->position[k] || !in->direction[k]) return fail(OPTK_EINVAL, "input position and direction are required");
  if (n == 0) return OPTK_OK;
  if (precision == OPTK_F32) {
    FRaysIn fi;
    for (int k = 0; k < 3; ++k) {
      fi.p[k] = (const float *)in->position[k];
      fi.d[k] = (const float *)in->direction[k];
    }
    fi.wavelength = (const float *)in->wavelength;
    fi.intensity = (const float *)in->intensity;
    fi.attenuation = (const float *)in->attenuation;
    fi.index = (const float *)in->index_refraction;
    fi.unvignetted = in->unvignetted;
    fi.wavelength_default = (float)in->wavelength_default;
    fi.intensity_default = (float)in->intensity_default;
    fi.attenuation_default = (float)in->attenuation_default;
    fi.index_default = (float)in->index_default;
    fi.unvignetted_default = in->unvignetted_default;
    FRaysOut fo, fa;
    fill_out32(fo, out);
    fill_out32(fa, accum);
    return fast_trace_rays(sys, fi, fo, accum ? &fa : nullptr, n, A.apply_post,
                           (flags & OPTK_TRACE_GENERIC) != 0, (hipStream_t)hip_stream);
  }
  RaysIn ri;
  for (int k = 0; k < 3; ++k) {
    ri.p[k] = (const double *)in->position[k];
    ri.d[k] = (const double *)in->direction[k];
  }
  ri.wavelength = (const double *)in->wavelength;
  ri.intensity = (const double *)in->intensity;
  ri.attenuation = (const double *)in->attenuation;
  ri.index = (const double *)in->index_refraction;
  ri.unvignetted = in->unvignetted;
  ri.wavelength_default = in->wavelength_default;
  ri.intensity_default = in->intensity_default;
  ri.attenuation_default = in->attenuation_default;
  ri.index_default = in->index_default;
  ri.unvignetted_default = in->unvignetted_default;
  const bool spec_ok = sys->signature >= 0 && !(flags & OPTK_TRACE_GENERIC) &&
                       n * (int64_t)sys->num_paths < ((int64_t)1 << 29);
  if (!accum && full_output(A.out) && spec_ok)
    hipLaunchKernelGGL(SPEC_RAYS64[sys->signature], dim3(spec_blocks(n)), dim3(TRACE_BLOCK), 0,
                       (hipStream_t)hip_stream, A, ri);
  else if (!accum && full_output(A.out))
    launch_trace<RaysIn>(trace_rays_f64<0, false, IO_FULL>,
                         trace_rays_f64<FEAT_TOROID | FEAT_RULINGS, false, IO_FULL>,
                         trace_rays_f64<FEAT_ALL, false, IO_FULL>, trace_rays_f64<FEAT_ALL, true>,
                         sys->features, false, n, (hipStream_t)hip_stream, A, ri, acc);
  else
    launch_trace<RaysIn>(trace_rays_f64<0, false>, trace_rays_f64<FEAT_TOROID | FEAT_RULINGS, false>,
                         trace_rays_f64<FEAT_ALL, false>, trace_rays_f64<FEAT_ALL, true>, sys->features,
                         accum != nullptr, n, (hipStream_t)hip_stream, A, ri, acc);
  OPTK_HIP_CHECK(hipGetLastError());
  return OPTK_OK;
}

int32_t optk_trace_grid(const optk_system *sys, int32_t precision, const optk_grid *in,
                        const optk_rays *out, const optk_rays *accum, int64_t n, int32_t flags,
                        void *hip_stream) {
  TraceArgs A;
  RaysOut acc;
  int32_t rc = prepare(sys, precision, out, accum, n, flags, A, acc);
  if (rc) return rc;
  if (!in) return fail(OPTK_EINVAL, "null input grid");
  if (n == 0) return OPTK_OK;
  if (precision == OPTK_F32) {
    FGridIn fg;
    for (int k = 0; k < 2; ++k) {
      fg.field[k] = (const float *)in->field[k];
      fg.pupil[k] = (const float *)in->pupil[k];
      fg.field_default[k] = (float)in->field_default[k];
      fg.pupil_default[k] = (float)in->pupil_default[k];
    }
    fg.wavelength = (const float *)in->wavelength;
    fg.intensity = (const float *)in->intensity;
    fg.wavelength_default = (float)in->wavelength_default;
    fg.intensity_default = (float)in->intensity_default;
    fg.object_at_infinity = in->object_at_infinity;
    FRaysOut fo, fa;
    fill_out32(fo, out);
    fill_out32(fa, accum);
    return fast_trace_grid(sys, fg, fo, accum ? &fa : nullptr, n, A.apply_post,
                           (flags & OPTK_TRACE_GENERIC) != 0, (hipStream_t)hip_stream);
  }
  GridIn g;
  for (int k = 0; k < 2; ++k) {
    g.field[k] = (const double *)in->field[k];
    g.pupil[k] = (const double *)in->pupil[k];
    g.field_default[k] = in->field_default[k];
    g.pupil_default[k] = in->pupil_default[k];
  }
  g.wavelength = (const double *)in->wavelength;
  g.intensity = (const double *)in->intensity;
  g.wavelength_default = in->wavelength_default;
  g.intensity_default = in->intensity_default;
  g.object_at_infinity = in->object_at_infinity;
  const bool spec_ok = sys->signature >= 0 && !(flags & OPTK_TRACE_GENERIC) &&
                       n * (int64_t)sys->num_paths < ((int64_t)1 << 29);
  if (!accum && sensor_layout(g, A.out) && spec_ok)
    hipLaunchKernelGGL(SPEC_GRID64[sys->signature], dim3(spec_blocks(n)), dim3(TRACE_BLOCK), 0,
                       (hipStream_t)hip_stream, A, g);
  else if (!accum && sensor_layout(g, A.out))
    launch_trace<GridIn>(trace_grid_f64<0, false, IO_SENSOR>,
                         trace_grid_f64<FEAT_TOROID | FEAT_RULINGS, false, IO_SENSOR>,
                         trace_grid_f64<FEAT_ALL, false, IO_SENSOR>, trace_grid_f64<FEAT_ALL, true>,
                         sys->features, false, n, (hipStream_t)hip_stream, A, g, acc);
  else
    launch_trace<GridIn>(trace_grid_f64<0, false>, trace_grid_f64<FEAT_TOROID | FEAT_RULINGS, false>,
                         trace_grid_f64<FEAT_ALL, false>, trace_grid_f64<FEAT_ALL, true>, sys->features,
                         accum != nullptr, n, (hipStream_t)hip_stream, A, g, acc);
  OPTK_HIP_CHECK(hipGetLastError());
  return OPTK_OK;
}

int32_t optk_eval(const optk_system *sys, int32_t op, const optk_eval_args *args, int64_t n,
                  void *hip_stream) {
  if (!sys || !args) return fail(OPTK_EINVAL, "null argument");
  if (op < OPTK_OP_SAG_VALUE || op > OPTK_OP_SNELL_SCALAR) return fail(OPTK_EINVAL, "unknown op");
  if (n < 0) return fail(OPTK_EINVAL, "negative count");
  bool needs_surface = !(op == OPTK_OP_SNELL || op == OPTK_OP_MULTILAYER || op == OPTK_OP_SNELL_SCALAR);
  if (needs_surface && (args->surface < 0 || args->surface >= sys->num_surfaces))
    return fail(OPTK_EINVAL, "surface index out of range");
  if (op == OPTK_OP_MULTILAYER && (args->stack < 0 || args->stack >= sys->num_stacks))
    return fail(OPTK_EINVAL, "stack index out of range");
  if ((op == OPTK_OP_INCIDENT || op == OPTK_OP_RULINGS_EFF)) {
    const optk_surface &s = sys->host_surfaces[args->surface];
    if (s.rulings_kind == OPTK_RULINGS_NONE) return fail(OPTK_EINVAL, "surface has no rulings");
    if (args->iparam < 0 || args->iparam >= (s.num_orders > 1 ? s.num_orders : 1))
      return fail(OPTK_EINVAL, "path index out of range");
  }
  static const int n_in[] = {3, 3, 8, 3, 8, 5, 6, 11, 10, 9, 2, 6};
  static const int n_out[] = {1, 3, 7, 0, 3, 12, 3, 3, 1, 1, 1, 2};
  for (int k = 0; k < n_in[op]; ++k)
    if (!args->in[k]) return fail(OPTK_EINVAL, "missing input array");
  for (int k = 0; k < n_out[op]; ++k)
    if (!args->out[k]) return fail(OPTK_EINVAL, "missing output array");
  if (op == OPTK_OP_APERTURE && !args->out_mask) return fail(OPTK_EINVAL, "missing output mask");
  if (n == 0) return OPTK_OK;
  EvalParams E;
  std::memset(&E, 0, sizeof(E));
  E.sys = view_of(sys);
  E.op = op;
  E.surface = args->surface;
  E.iparam = args->iparam;
  E.stack = args->stack;
  E.n = n;
  for (int k = 0; k < 12; ++k) {
    E.in[k] = args->in[k];
    E.out[k] = args->out[k];
  }
  E.out_mask = args->out_mask;
  hipLaunchKernelGGL(eval_f64, grid_for(n), dim3(256), 0, (hipStream_t)hip_stream, E);
  OPTK_HIP_CHECK(hipGetLastError());
  return OPTK_OK;
}

int32_t optk_stop_newton(const optk_system *sys, const optk_system *first,
                         const optk_stop_args *args, int64_t n, void *hip_stream) {
  if (!sys || !args) return fail(OPTK_EINVAL, "null argument");
  if (n < 0) return fail(OPTK_EINVAL, "negative count");
  const optk_stop_args &a = *args;
  if (a.free_component != OPTK_STOP_FREE_DIRECTION && a.free_component != OPTK_STOP_FREE_POSITION)
    return fail(OPTK_EINVAL, "free_component must be OPTK_STOP_FREE_DIRECTION or _POSITION");
  if (a.target_component != OPTK_STOP_TARGET_POSITION && a.target_component != OPTK_STOP_TARGET_DIRECTION)
    return fail(OPTK_EINVAL, "target_component must be OPTK_STOP_TARGET_POSITION or _DIRECTION");
  if (a.free_component == OPTK_STOP_FREE_POSITION && (!first || first->num_surfaces < 1))
    return fail(OPTK_EINVAL, "a free position needs the first surface's sag");
  if (a.max_iter < 0 || a.max_iter > 10000) return fail(OPTK_EINVAL, "max_iter out of range");
  if (!(a.dx > 0.0) || !(a.tol > 0.0)) return fail(OPTK_EINVAL, "dx and tol must be positive");
  for (int k = 0; k < 3; ++k)
    if (!a.position[k] || !a.direction[k]) return fail(OPTK_EINVAL, "missing ray array");
  if (!a.wavelength || !a.target[0] || !a.target[1] || !a.x[0] || !a.x[1] || !a.iterations)
    return fail(OPTK_EINVAL, "missing array");
  if (n == 0) return OPTK_OK;
  StopParams P;
  std::memset(&P, 0, sizeof(P));
  P.T.spec = sys->spec;
  P.T.surfaces = sys->surfaces;
  P.T.num_surfaces = sys->num_surfaces;
  P.T.apply_post = 1;
  P.T.n = n;
  if (first && first->num_surfaces > 0) P.first = first->host_surfaces[0];
  P.a = a;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(stop_newton_f64<FEAT_ALL>, dim3((unsigned)blocks), dim3(256),
                     staged_bytes(sys->num_surfaces), (hipStream_t)hip_stream, P);
  OPTK_HIP_CHECK(hipGetLastError());
  return OPTK_OK;
}

int32_t optk_pupil_reduce(const optk_pupil_args *args, int64_t segments, void *hip_stream) {
  if (!args) return fail(OPTK_EINVAL, "null argument");
  if (segments < 0 || args->per_segment < 0) return fail(OPTK_EINVAL, "negative count");
  if (!args->unvignetted) return fail(OPTK_EINVAL, "missing unvignetted mask");
  if ((args->sum_x || args->all_x) && !args->x) return fail(OPTK_EINVAL, "missing x");
  if ((args->sum_y || args->all_y) && !args->y) return fail(OPTK_EINVAL, "missing y");
  if (segments == 0) return OPTK_OK;
  const int64_t threads = segments * 64;
  const int64_t blocks = (threads + 255) / 256;
  if (blocks > INT32_MAX) return fail(OPTK_EINVAL, "too many segments");
  hipLaunchKernelGGL(pupil_reduce_kernel, dim3((unsigned)blocks), dim3(256), 0,
                     (hipStream_t)hip_stream, *args, segments);
  OPTK_HIP_CHECK(hipGetLastError());
  return OPTK_OK;
}

int32_t optk_denormalize(const optk_denormalize_args *args, int64_t n, void *hip_stream) {
  if (!args) return fail(OPTK_EINVAL, "null argument");
  if (n < 0) return fail(OPTK_EINVAL, "negative count");
  if (args->n_table < 1) return fail(OPTK_EINVAL, "empty extent table");
  int ncomp = 0;
  while (ncomp < 4 && args->in[ncomp]) ++ncomp;
  for (int c = 0; c < ncomp; ++c)
    if (!args->out[c] || !args->lo[c] || !args->ptp[c])
      return fail(OPTK_EINVAL, "missing output or extent array");
  if (n == 0 || ncomp == 0) return OPTK_OK;
  if ((n + 255) / 256 > INT32_MAX) return fail(OPTK_EINVAL, "too many values");
  hipLaunchKernelGGL(denormalize_kernel, grid_for(n), dim3(256), 0, (hipStream_t)hip_stream, *args,
                     ncomp, n);
  OPTK_HIP_CHECK(hipGetLastError());
  return OPTK_OK;
}

int32_t optk_expose(const optk_expose_args *args, int64_t n, void *hip_stream) {
  if (!args) return fail(OPTK_EINVAL, "null argument");
  if (n < 0) return fail(OPTK_EINVAL, "negative count");
  if (!args->photons || !args->electrons) return fail(OPTK_EINVAL, "missing image");
  if (!(args->scale >= 0.0) || !(args->read_noise >= 0.0))
    return fail(OPTK_EINVAL, "scale and read_noise must be non-negative");
  if (n == 0) return OPTK_OK;
  hipLaunchKernelGGL(expose_kernel, grid_for(n), dim3(256), 0, (hipStream_t)hip_stream, *args, n);
  OPTK_HIP_CHECK(hipGetLastError());
  return OPTK_OK;
}


// Workspace of the tiled binning (sections 256-byte aligned); fills T's
// pointers when ws is given.  0 when the image has too many tiles or the
// rays are too many for 32-bit slots.
static uint64_t tiled_layout(int nw, int nx, int ny, int64_t n, bool f32, void *ws,
                             TiledParams *T) {
  const int64_t tx = (nx + CT_SIDE - 1) / CT_SIDE, ty = (ny + CT_SIDE - 1) / CT_SIDE;
  const int64_t tiles = (int64_t)nw * tx * ty;
  if (tiles > CT_MAX_TILES || n <= 0 || n > INT32_MAX) return 0;
  const int64_t blocks = (n + CT_CHUNK - 1) / CT_CHUNK;
  auto up = [](uint64_t b) { return (b + 255) & ~(uint64_t)255; };
  const uint64_t s_counts = up((uint64_t)blocks * tiles * 4);
  const uint64_t s_base = up((uint64_t)(tiles + 1) * 4);
  const uint64_t s_keys = up((uint64_t)n * 4);
  const uint64_t s_pix = f32 ? 0 : up((uint64_t)n * 2);
  const uint64_t s_val = up((uint64_t)n * 16);
  const uint64_t total = s_counts + s_base + s_keys + s_pix + s_val;
  if (T) {
    T->tiles_x = (int32_t)tx;
    T->tiles_y = (int32_t)ty;
    T->num_tiles = (int32_t)tiles;
    T->blocks = (int32_t)blocks;
    char *p = (char *)ws;
    T->counts = (int32_t *)p;
    p += s_counts;
    T->base = (int32_t *)p;
    p += s_base;
    T->keys = (int32_t *)p;
    p += s_keys;
    T->sorted_pix = f32 ? nullptr : (uint16_t *)p;
    p += s_pix;
    T->sorted_ab = f32 ? nullptr : (double2 *)p;
    T->sorted32 = f32 ? (RecordF32 *)p : nullptr;
  }
  return total;
}

}  // extern "C"

// Sensor binning with an explicit workspace (ws may be null: the one-pass
// kernel with float64 global atomics).  Shared by optk_collect (the
// caller's workspace) and optk_collect_noalloc (the library's scratch).
// Arguments of a binning call checked before any device work (OPTK_OK or
// an error).
static int32_t collect_validate(const optk_collect_args *a, int64_t n) {
  if (!a) return fail(OPTK_EINVAL, "null argument");
  if (n < 0) return fail(OPTK_EINVAL, "negative count");
  if (a->n_wavelength < 1 || a->nx < 1 || a->ny < 1) return fail(OPTK_EINVAL, "empty histogram");
  if (!a->position[0] || !a->position[1] || !a->direction_z || !a->wavelength || !a->intensity ||
      !a->wavelength_edges || !a->image || !a->moment)
    return fail(OPTK_EINVAL, "missing array");
  if (!(a->x_hi > a->x_lo) || !(a->y_hi > a->y_lo)) return fail(OPTK_EINVAL, "empty pixel range");
  if (a->precision != OPTK_F64 && a->precision != OPTK_F32) return fail(OPTK_EINVAL, "bad precision");
  return OPTK_OK;
}

static int32_t collect_launch(const optk_collect_args *a, int64_t n, void *ws, uint64_t ws_bytes,
                              void *hip_stream) {
  const int32_t rc = collect_validate(a, n);
  if (rc) return rc;
  if (n == 0) return OPTK_OK;
  CollectParams C;
  std::memset(&C, 0, sizeof(C));
  if (a->precision == OPTK_F64) {
    C.px = (const double *)a->position[0];
    C.py = (const double *)a->position[1];
    C.dz = (const double *)a->direction_z;
    C.wavelength = (const double *)a->wavelength;
    C.intensity = (const double *)a->intensity;
  } else {
    C.pxf = (const float *)a->position[0];
    C.pyf = (const float *)a->position[1];
    C.dzf = (const float *)a->direction_z;
    C.wavelengthf = (const float *)a->wavelength;
    C.intensityf = (const float *)a->intensity;
  }
  C.unvignetted = a->unvignetted;
  C.wedges = a->wavelength_edges;
  C.nw = a->n_wavelength;
  C.nx = a->nx;
  C.ny = a->ny;
  C.x_lo = a->x_lo;
  C.x_hi = a->x_hi;
  C.y_lo = a->y_lo;
  C.y_hi = a->y_hi;
  C.image = a->image;
  C.moment = a->moment;
  C.n = n;
  // numpy.linspace's step, as the edges are formed (e_j = j * step + lo)
  C.x_step = (C.x_hi - C.x_lo) / (double)C.nx;
  C.y_step = (C.y_hi - C.y_lo) / (double)C.ny;
  C.x_inv = 1.0 / C.x_step;
  C.y_inv = 1.0 / C.y_step;
  const bool f32 = a->precision == OPTK_F32;
  hipStream_t s = (hipStream_t)hip_stream;
  // the first-guess bin search is exact while num * (|lo| + |hi|) / (hi - lo)
  // < 2^48; a range beyond (narrow for its distance from the origin) is
  // binned by the one-pass kernel with a binary search over the edges
  auto narrow = [](double lo, double hi, int num) {
    return !((double)num * (fabs(lo) + fabs(hi)) < 281474976710656.0 * (hi - lo));
  };
  if (narrow(C.x_lo, C.x_hi, C.nx) || narrow(C.y_lo, C.y_hi, C.ny)) {
    hipLaunchKernelGGL(collect_kernel<true>, grid_for(n), dim3(256), 0, s, C);
    OPTK_HIP_CHECK(hipGetLastError());
    return OPTK_OK;
  }
  TiledParams T;
  const uint64_t need = tiled_layout(a->n_wavelength, a->nx, a->ny, n, f32, ws, &T);
  if (need == 0 || !ws || ws_bytes < need) {
    hipLaunchKernelGGL(collect_kernel<false>, grid_for(n), dim3(256), 0, s, C);
    OPTK_HIP_CHECK(hipGetLastError());
    return OPTK_OK;
  }
  T.C = C;
  const size_t lds = sizeof(int32_t) * (size_t)T.num_tiles;
  if (f32)
    hipLaunchKernelGGL(collect_count<true>, dim3((unsigned)T.blocks), dim3(CT_THREADS), lds, s, T);
  else
    hipLaunchKernelGGL(collect_count<false>, dim3((unsigned)T.blocks), dim3(CT_THREADS), lds, s, T);
  hipLaunchKernelGGL(collect_offsets, dim3((unsigned)((T.num_tiles + 3) / 4)), dim3(256), 0, s, T);
  hipLaunchKernelGGL(collect_base, dim3(1), dim3(1024), 0, s, T);
  if (T.num_tiles <= CT_STAGED_TILES) {
    if (f32) {
      hipLaunchKernelGGL(collect_scatter_staged<true>, dim3((unsigned)T.blocks), dim3(CT_THREADS),
                         3 * lds, s, T);
      hipLaunchKernelGGL(collect_tiles<true>, dim3((unsigned)(T.num_tiles * CT_SPLIT)), dim3(256), 0, s, T);
    } else {
      hipLaunchKernelGGL(collect_scatter_staged<false>, dim3((unsigned)T.blocks), dim3(CT_THREADS),
                         3 * lds, s, T);
      hipLaunchKernelGGL(collect_tiles<false>, dim3((unsigned)(T.num_tiles * CT_SPLIT)), dim3(256), 0, s, T);
    }
  } else if (f32) {
    hipLaunchKernelGGL(collect_scatter<true>, dim3((unsigned)T.blocks), dim3(CT_THREADS), lds, s, T);
    hipLaunchKernelGGL(collect_tiles<true>, dim3((unsigned)(T.num_tiles * CT_SPLIT)), dim3(256), 0, s, T);
  } else {
    hipLaunchKernelGGL(collect_scatter<false>, dim3((unsigned)T.blocks), dim3(CT_THREADS), lds, s, T);
    hipLaunchKernelGGL(collect_tiles<false>, dim3((unsigned)(T.num_tiles * CT_SPLIT)), dim3(256), 0, s, T);
  }
  OPTK_HIP_CHECK(hipGetLastError());
  return OPTK_OK;
}

// The no-allocation binning's workspace: one array per GPU, held by the
// code object until the process ends.  Host code never passes `scratch`
// itself to a kernel or a copy (on the host the name is a stand-in, not a
// device address): it uses the address hipGetSymbolAddress returns for the
// current device, looked up once per device and kept in scratch_ptr.
static __device__ __align__(256) unsigned char scratch[OPTK_COLLECT_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
void *scratch_ptr[kMaxDevices] = {};
}  // namespace

// the current device's address of `scratch` (looked up on its first use)
static int32_t scratch_address(void **out) {
  int dev = -1;
  OPTK_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(OPTK_EDEVICE, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(OPTK_EDEVICE, std::string("hipGetSymbolAddress(scratch) failed: ") +
                                    hipGetErrorString(e));
    scratch_ptr[dev] = p;
  }
  *out = scratch_ptr[dev];
  return OPTK_OK;
}

extern "C" {

int32_t optk_collect(const optk_collect_args *a, int64_t n, void *hip_stream) {
  if (!a) return fail(OPTK_EINVAL, "null argument");
  return collect_launch(a, n, a->workspace, a->workspace_bytes, hip_stream);
}

int32_t optk_collect_noalloc(const optk_collect_args *a, int64_t n, void *hip_stream) {
  const int32_t bad = collect_validate(a, n);
  if (bad) return bad;
  uint64_t need = 0;
  if (n > 0) {
    // more rays than 32-bit slots hold: the tiled path does not apply, and
    // the one-pass path would run instead; refuse, as for any size whose
    // tiled workspace exceeds the scratch
    if (n > INT32_MAX)
      return fail(OPTK_EINVAL, "optk_collect_noalloc: at most 2^31 - 1 rays per call");
    need = tiled_layout(a->n_wavelength, a->nx, a->ny, n, a->precision == OPTK_F32, nullptr,
                        nullptr);
    if (need > OPTK_COLLECT_SCRATCH_BYTES)
      return fail(OPTK_EINVAL,
                  "optk_collect_noalloc: " + std::to_string(n) + " rays need " +
                      std::to_string(need) + " bytes of workspace, more than the " +
                      std::to_string((unsigned long long)OPTK_COLLECT_SCRATCH_BYTES) +
                      "-byte scratch; split the rays over several calls or use optk_collect "
                      "with a workspace");
  }
  void *ws = nullptr;
  if (need > 0) {
    const int32_t rc = scratch_address(&ws);
    if (rc) return rc;
  }
  return collect_launch(a, n, ws, need, hip_stream);
}

uint64_t optk_collect_workspace(const optk_collect_args *a, int64_t n) {
  if (!a || n <= 0 || a->n_wavelength < 1 || a->nx < 1 || a->ny < 1) return 0;
  return tiled_layout(a->n_wavelength, a->nx, a->ny, n, a->precision == OPTK_F32, nullptr, nullptr);
}

}  // extern "C"
