// optk_fast.hip -- gfx950 kernels of the float32 form of the trace
// (OPTK_F32; physics in optk_fast.h).  Compiled with -ffp-contract=fast:
// unlike the float64 form, this form does not reproduce numpy's rounding
// sequence, so multiply-adds may fuse.

#include "optk_batch.h"
#include "optk_signatures.h"

using namespace optk;

namespace {

struct FastArgs {
  const FastSpec *spec;
  const FastSurface *fast;
  const optk_surface *surfaces;
  StackView sv;
  const float *tables32;
  int32_t num_surfaces;
  int32_t num_paths;
  int32_t entry;       // 0: rays in the system frame, 1: grid (pre-maps)
  int32_t apply_post;  // 1: leave in the post (sensor) frame
  int64_t n;
  FRaysOut out;
};

static_assert(sizeof(FastSurface) % 8 == 0 && sizeof(FastSpec) % 8 == 0 &&
                  sizeof(optk_surface) % 8 == 0,
              "records are staged as words");

__device__ __forceinline__ float ldf(const float *p, int64_t i, float dflt) {
  return p ? p[i] : dflt;
}

__device__ __forceinline__ FRay load_fray(const FRaysIn &in, int64_t i) {
  FRay r;
  r.px = in.p[0][i];
  r.py = in.p[1][i];
  r.pz = in.p[2][i];
  r.dx = in.d[0][i];
  r.dy = in.d[1][i];
  r.dz = in.d[2][i];
  r.wavelength = ldf(in.wavelength, i, in.wavelength_default);
  r.intensity = ldf(in.intensity, i, in.intensity_default);
  r.attenuation = ldf(in.attenuation, i, in.attenuation_default);
  r.index = ldf(in.index, i, in.index_default);
  r.unvignetted = in.unvignetted ? (in.unvignetted[i] != 0) : (in.unvignetted_default != 0);
  return r;
}

__device__ __forceinline__ FRay load_fray(const FGridIn &g, int64_t i) {
  return fast_grid_ray(ldf(g.field[0], i, g.field_default[0]), ldf(g.field[1], i, g.field_default[1]),
                       ldf(g.pupil[0], i, g.pupil_default[0]), ldf(g.pupil[1], i, g.pupil_default[1]),
                       ldf(g.wavelength, i, g.wavelength_default),
                       ldf(g.intensity, i, g.intensity_default), g.object_at_infinity);
}

// a streaming store: the trace writes each output element once and never
// reads it back, so the store is marked nontemporal and does not keep the
// line in the caches (the Newtonian sensor-frame kernels measured 8 %
// faster in float32 and 2-3 % in float64 with it; nontemporal loads of the
// inputs gained nothing)
template <typename T> __device__ __forceinline__ void put(T &dst, T v) { __builtin_nontemporal_store(v, &dst); }

__device__ __forceinline__ void store_f(const FRaysOut &o, int64_t j, const FRay &r) {
  if (o.p[0]) put<float>(o.p[0][j], r.px);
  if (o.p[1]) put<float>(o.p[1][j], r.py);
  if (o.p[2]) put<float>(o.p[2][j], r.pz);
  if (o.d[0]) put<float>(o.d[0][j], r.dx);
  if (o.d[1]) put<float>(o.d[1][j], r.dy);
  if (o.d[2]) put<float>(o.d[2][j], r.dz);
  if (o.wavelength) put<float>(o.wavelength[j], r.wavelength);
  if (o.intensity) put<float>(o.intensity[j], r.intensity);
  if (o.attenuation) put<float>(o.attenuation[j], r.attenuation);
  if (o.index) put<float>(o.index[j], r.index);
  if (o.unvignetted) put<uint8_t>(o.unvignetted[j], r.unvignetted ? 1 : 0);
}

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

bool full_output(const FRaysOut &o) {
  return o.p[0] && o.p[1] && o.p[2] && o.d[0] && o.d[1] && o.d[2] && o.wavelength && o.intensity &&
         o.attenuation && o.index && o.unvignetted;
}

bool sensor_layout(const FGridIn &g, const FRaysOut &o) {
  return g.field[0] && g.field[1] && g.pupil[0] && g.pupil[1] && g.wavelength && !g.intensity &&
         o.p[0] && o.p[1] && o.p[2] && o.d[0] && o.d[1] && o.d[2] && o.intensity &&
         o.unvignetted && !o.wavelength && !o.attenuation && !o.index;
}

template <int IO> __device__ __forceinline__ FRay load_ray(const FRaysIn &in, int64_t i) {
  return load_fray(in, i);
}

template <int IO> __device__ __forceinline__ FRay load_ray(const FGridIn &g, int64_t i) {
  if (IO == IO_SENSOR)
    return fast_grid_ray(g.field[0][i], g.field[1][i], g.pupil[0][i], g.pupil[1][i],
                         g.wavelength[i], g.intensity_default, g.object_at_infinity);
  return load_fray(g, i);
}

template <int IO> __device__ __forceinline__ void store_ray(const FRaysOut &o, int64_t j, const FRay &r) {
  if (IO == IO_FULL) {
    put<float>(o.p[0][j], r.px);
    put<float>(o.p[1][j], r.py);
    put<float>(o.p[2][j], r.pz);
    put<float>(o.d[0][j], r.dx);
    put<float>(o.d[1][j], r.dy);
    put<float>(o.d[2][j], r.dz);
    put<float>(o.wavelength[j], r.wavelength);
    put<float>(o.intensity[j], r.intensity);
    put<float>(o.attenuation[j], r.attenuation);
    put<float>(o.index[j], r.index);
    put<uint8_t>(o.unvignetted[j], r.unvignetted ? 1 : 0);
    return;
  }
  if (IO == IO_SENSOR) {
    put<float>(o.p[0][j], r.px);
    put<float>(o.p[1][j], r.py);
    put<float>(o.p[2][j], r.pz);
    put<float>(o.d[0][j], r.dx);
    put<float>(o.d[1][j], r.dy);
    put<float>(o.d[2][j], r.dz);
    put<float>(o.intensity[j], r.intensity);
    put<uint8_t>(o.unvignetted[j], r.unvignetted ? 1 : 0);
    return;
  }
  store_f(o, j, r);
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
template <int IO> __device__ __forceinline__ FRay load_ray32(const FGridIn &g, uint32_t i) {
  if (IO == IO_SENSOR)
    return fast_grid_ray(at32(g.field[0], i), at32(g.field[1], i), at32(g.pupil[0], i), at32(g.pupil[1], i),
                         at32(g.wavelength, i), g.intensity_default, g.object_at_infinity);
  return load_ray<IO>(g, (int64_t)i);
}
template <int IO> __device__ __forceinline__ FRay load_ray32(const FRaysIn &in, uint32_t i) {
  return load_ray<IO>(in, (int64_t)i);
}

template <int IO> __device__ __forceinline__ void store_ray32(const FRaysOut &o, uint32_t j, const FRay &r) {
  if (IO == IO_SENSOR) {
    // the offset is formed here (an empty asm stops it being hoisted into
    // an earlier block as a 64-bit address, which the stores could then not
    // take as a scalar base plus a 32-bit offset)
    asm volatile("" : "+v"(j));
    put<float>(at32(o.p[0], j), r.px);
    put<float>(at32(o.p[1], j), r.py);
    put<float>(at32(o.p[2], j), r.pz);
    put<float>(at32(o.d[0], j), r.dx);
    put<float>(at32(o.d[1], j), r.dy);
    put<float>(at32(o.d[2], j), r.dz);
    put<float>(at32(o.intensity, j), r.intensity);
    put<uint8_t>(at32(o.unvignetted, j), r.unvignetted ? 1 : 0);
    return;
  }
  store_ray<IO>(o, (int64_t)j, r);
}

size_t fast_staged_bytes(int num_surfaces) {
  return sizeof(FastSpec) + (sizeof(FastSurface) + sizeof(optk_surface)) * (size_t)num_surfaces;
}

// occupancy targets (waves per SIMD): the geometric kernel at 6 (80 VGPRs;
// measured faster than 5 or 8 on the Newtonian workload), the rulings /
// toroid kernel at 4, the multilayer kernels at 2 (VALU-bound, register-heavy)
template <int F> struct FastWaves {
  static constexpr int value = F == 0 ? 6 : ((F & FEAT_MULTILAYER) ? 2 : 4);
};
// the specialised kernels: 3 waves for the multilayer layout (168 VGPRs)
// measured 2 % faster on the spectrograph than 2 (256 VGPRs), 4 (128
// VGPRs) 16 % slower
template <int F> struct FastSpecWaves {
  static constexpr int value = F == 0 ? 6 : ((F & FEAT_MULTILAYER) ? 3 : 4);
};
// the batched kernels keep 4 / 2 (at 6 the geometric one spills)
template <int F> struct BatchWaves {
  static constexpr int value = (F & FEAT_MULTILAYER) ? 2 : 4;
};

// The system's records are read through the constant address space: every
// lane of a wave reads the same record at the same time, so the loads are
// scalar-cache reads whose values land in SGPRs, which the float64 VALU
// instructions take as operands directly (no LDS staging, no VGPRs holding
// matrix elements, branch conditions already uniform).
typedef __attribute__((address_space(4))) const FastSpec ConstFastSpec;
typedef __attribute__((address_space(4))) const FastSurface ConstFastSurface;
typedef __attribute__((address_space(4))) const optk_surface ConstSurface;

template <int F, bool ACCUM, int IO, typename In>
__device__ __forceinline__ void fast_body(const FastArgs &A, const In &in, const FRaysOut &accum) {
  ConstFastSpec *c_spec = (ConstFastSpec *)A.spec;
  ConstFastSurface *c_fast = (ConstFastSurface *)A.fast;
  ConstSurface *c_surf = (ConstSurface *)A.surfaces;
  FastView v;
  v.surfaces = A.surfaces;
  v.fast = A.fast;
  v.sv = A.sv;
  v.tables32 = A.tables32;
  const int num_paths = (F & FEAT_RULINGS) ? A.num_paths : 1;
  const int num_surfaces = A.num_surfaces;
  const int64_t n = A.n;
  // grid entry without per-surface accumulation: a pass-through surface 0
  // is elided (FastSpec::skip_first)
  const bool skip = A.entry == 1 && !ACCUM && c_spec->skip_first != 0;
  const int k0 = skip ? 1 : 0;
  // the launch block is TRACE_BLOCK threads: a compile-time stride keeps
  // it in scalar registers (blockDim.x would be a per-lane load)
  const int64_t stride = (int64_t)gridDim.x * TRACE_BLOCK;
  for (int64_t i = (int64_t)blockIdx.x * TRACE_BLOCK + threadIdx.x; i < n; i += stride) {
    FRay r0 = load_ray<IO>(in, i);
    if (skip) {
      // surface 0 is a pass-through on whose plane the grid rays start; a ray
      // parallel to it meets it at t = -0/0 = NaN
      const bool parallel = r0.dz == 0.0;
      fast_xf(*(const optk_affine *)&c_spec->first_skip, r0);
      if (parallel) {
        r0.px = r0.py = r0.pz = NAN;
        r0.intensity = NAN;
      }
    } else {
      fast_xf(*(const optk_affine *)&c_spec->first_hop[A.entry], r0);
    }
    for (int path = 0; path < num_paths; ++path) {
      FRay r = r0;
      for (int k = k0; k < num_surfaces; ++k) {
        const FastSurface &fk = *(const FastSurface *)&c_fast[k];
        const optk_surface &sk = *(const optk_surface *)&c_surf[k];
        if (k > k0) {
          fast_xf(fk.hop, r);
        }
        fast_propagate<F>(fk, sk, v, path, r);
        if (ACCUM) {
          FRay g = r;
          fast_xf(fk.to_global, g);
          store_f(accum, ((int64_t)k * num_paths + path) * n + i, g);
        }
      }
      const optk_affine &last = *(const optk_affine *)&c_spec->last[A.apply_post];
      fast_xf(last, r);
      store_ray<IO>(A.out, (int64_t)path * n + i, r);
    }
  }
}

template <int F, bool ACCUM, int IO = IO_ANY>
__global__ __launch_bounds__(TRACE_BLOCK) __attribute__((amdgpu_waves_per_eu(FastWaves<F>::value)))
void trace_rays_f32(const FastArgs A, const FRaysIn in, const FRaysOut accum) {
  fast_body<F, ACCUM, IO>(A, in, accum);
}

template <int F, bool ACCUM, int IO = IO_ANY>
__global__ __launch_bounds__(TRACE_BLOCK) __attribute__((amdgpu_waves_per_eu(FastWaves<F>::value)))
void trace_grid_f32(const FastArgs A, const FGridIn g, const FRaysOut accum) {
  fast_body<F, ACCUM, IO>(A, g, accum);
}

FastArgs make_args(const optk_system *sys, const FRaysOut &out, int64_t n, int32_t entry,
                   int32_t apply_post) {
  FastArgs A;
  memset(&A, 0, sizeof(A));
  A.spec = sys->fast_spec;
  A.fast = sys->fast;
  A.surfaces = sys->surfaces;
  A.sv.grids = sys->grids;
  A.sv.grid_lo = sys->grid_lo;
  A.sv.layers = sys->layers;
  A.sv.stacks = sys->stacks;
  A.sv.chemicals = sys->chemicals;
  A.sv.tables = sys->tables;
  A.tables32 = sys->tables32;
  A.num_surfaces = sys->num_surfaces;
  A.num_paths = sys->num_paths;
  A.entry = entry;
  A.apply_post = apply_post ? 1 : 0;
  A.n = n;
  A.out = out;
  return A;
}

// kernels of one entry point: by feature set (geometric, toroid + rulings,
// Kernels specialised on a system's surface kinds (optk_signatures.h), one
// ray per thread, the surfaces unrolled in order with their kinds fixed at
// compile time; the arithmetic is that of fast_body.
template <int F, uint32_t K>
__device__ __forceinline__ void fast_spec_step(int k, int k0, ConstFastSurface *c_fast,
                                               ConstSurface *c_surf, const FastView &v, int path,
                                               FRay &r) {
  if (k < k0) return;
  const FastSurface &fk = *(const FastSurface *)&c_fast[k];
  const optk_surface &sk = *(const optk_surface *)&c_surf[k];
  if (k > k0) fast_xf(fk.hop, r);
  fast_propagate<F, K>(fk, sk, v, path, r);
}

template <typename In, int F, int IO, uint32_t... K>
__device__ __forceinline__ void fast_spec_body(const FastArgs &A, const In &in) {
  // the host launches these kernels with n * paths < 2^29: indices and
  // every element offset fit 32 bits
  const uint32_t i32 = blockIdx.x * (uint32_t)TRACE_BLOCK + threadIdx.x, n32 = (uint32_t)A.n;
  if (i32 >= n32) return;
  ConstFastSpec *c_spec = (ConstFastSpec *)A.spec;
  ConstFastSurface *c_fast = (ConstFastSurface *)A.fast;
  ConstSurface *c_surf = (ConstSurface *)A.surfaces;
  FastView v;
  v.surfaces = A.surfaces;
  v.fast = A.fast;
  v.sv = A.sv;
  v.tables32 = A.tables32;
  const int num_paths = (F & FEAT_RULINGS) ? A.num_paths : 1;
  const bool skip = A.entry == 1 && c_spec->skip_first != 0;
  const int k0 = skip ? 1 : 0;
  FRay r0 = load_ray32<IO>(in, i32);
  if (skip) {
    const bool parallel = r0.dz == 0.0;
    fast_xf(*(const optk_affine *)&c_spec->first_skip, r0);
    if (parallel) {
      r0.px = r0.py = r0.pz = NAN;
      r0.intensity = NAN;
    }
  } else {
    fast_xf(*(const optk_affine *)&c_spec->first_hop[A.entry], r0);
  }
  for (int path = 0; path < num_paths; ++path) {
    FRay r = r0;
    int k = 0;
    (fast_spec_step<F, K>(k++, k0, c_fast, c_surf, v, path, r), ...);
    const optk_affine &last = *(const optk_affine *)&c_spec->last[A.apply_post];
    fast_xf(last, r);
    store_ray32<IO>(A.out, (uint32_t)path * n32 + i32, r);
  }
}

template <int F, int IO, uint32_t... K>
__global__ __launch_bounds__(TRACE_BLOCK) __attribute__((amdgpu_waves_per_eu(FastSpecWaves<F>::value)))
void trace_grid_f32_spec(const FastArgs A, const FGridIn g) {
  fast_spec_body<FGridIn, F, IO, K...>(A, g);
}

template <int F, int IO, uint32_t... K>
__global__ __launch_bounds__(TRACE_BLOCK) __attribute__((amdgpu_waves_per_eu(FastSpecWaves<F>::value)))
void trace_rays_f32_spec(const FastArgs A, const FRaysIn in) {
  fast_spec_body<FRaysIn, F, IO, K...>(A, in);
}

typedef void (*SpecGrid32)(FastArgs, FGridIn);
typedef void (*SpecRays32)(FastArgs, FRaysIn);
const SpecGrid32 SPEC_GRID32[NUM_SIGNATURES] = {
    trace_grid_f32_spec<SIGNATURES[SIG_NEWTONIAN].features, IO_SENSOR, OPTK_SIGNATURE_NEWTONIAN>,
    trace_grid_f32_spec<SIGNATURES[SIG_SPECTROGRAPH].features, IO_SENSOR, OPTK_SIGNATURE_SPECTROGRAPH>,
    trace_grid_f32_spec<SIGNATURES[SIG_SPECTROGRAPH_TILTED].features, IO_SENSOR,
                        OPTK_SIGNATURE_SPECTROGRAPH_TILTED>,
};
const SpecRays32 SPEC_RAYS32[NUM_SIGNATURES] = {
    trace_rays_f32_spec<SIGNATURES[SIG_NEWTONIAN].features, IO_FULL, OPTK_SIGNATURE_NEWTONIAN>,
    trace_rays_f32_spec<SIGNATURES[SIG_SPECTROGRAPH].features, IO_FULL, OPTK_SIGNATURE_SPECTROGRAPH>,
    trace_rays_f32_spec<SIGNATURES[SIG_SPECTROGRAPH_TILTED].features, IO_FULL,
                        OPTK_SIGNATURE_SPECTROGRAPH_TILTED>,
};

// kernels of one entry point: by feature set (geometric, toroid + rulings,
// all; all with the per-surface accumulation), for either I/O layout
template <typename In> struct KernelSet {
  typedef void (*Kernel)(FastArgs, In, FRaysOut);
  Kernel any[4];
  Kernel special[3];  // the entry point's fixed-layout form (IO_SENSOR for the
                      // grid, IO_FULL for rays), by feature set
};

template <typename In>
int32_t launch(const KernelSet<In> &ks, bool special, const optk_system *sys, const In &in,
               const FRaysOut &out, const FRaysOut *accum, int64_t n, int32_t entry,
               int32_t apply_post, void (*spec)(FastArgs, In), hipStream_t stream) {
  if (n == 0) return OPTK_OK;
  FastArgs A = make_args(sys, out, n, entry, apply_post);
  // (the specialised kernels address their arrays with 32-bit offsets)
  if (!accum && special && spec && n * (int64_t)sys->num_paths < ((int64_t)1 << 29)) {
    hipLaunchKernelGGL(spec, dim3((unsigned)((n + TRACE_BLOCK - 1) / TRACE_BLOCK)),
                       dim3(TRACE_BLOCK), 0, stream, A, in);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(OPTK_EDEVICE, std::string("float32 trace launch: ") + hipGetErrorString(e));
    return OPTK_OK;
  }
  FRaysOut acc;
  memset(&acc, 0, sizeof(acc));
  if (accum) acc = *accum;
  const int f = sys->features;
  const int which = f == 0 ? 0 : (f == (FEAT_TOROID | FEAT_RULINGS) ? 1 : 2);
  typename KernelSet<In>::Kernel kern = accum ? ks.any[3] : ks.any[which];
  if (!accum && special) kern = ks.special[which];
  // whole rounds of resident blocks; each block loops over its share
  const int64_t blocks = resident_grid((const void *)kern, TRACE_BLOCK, n);
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(TRACE_BLOCK), 0, stream, A, in, acc);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(OPTK_EDEVICE, std::string("float32 trace launch: ") + hipGetErrorString(e));
  return OPTK_OK;
}

// batched persistent trace, float32 form ------------------------------------

template <int F> struct Tracer32 {
  const BatchArgs &A;
  const FGridIn &in;
  const FRaysOut &out;
  FastSpec *s_spec;
  FastSurface *s_fast;
  optk_surface *s_surf;

  __device__ int num_paths() const { return (F & FEAT_RULINGS) ? A.num_paths : 1; }

  __device__ __forceinline__ void stage(int instance) {
    stage_words(s_spec, A.fast_specs + instance, sizeof(FastSpec));
    stage_words(s_fast, A.fast + (int64_t)instance * A.num_surfaces,
                sizeof(FastSurface) * (size_t)A.num_surfaces);
    stage_words(s_surf, A.surfaces + (int64_t)instance * A.num_surfaces,
                sizeof(optk_surface) * (size_t)A.num_surfaces);
  }

  __device__ __forceinline__ void trace(int64_t i, int path, double &x, double &y, double &inten,
                                        bool &unv) {
    FastView v;
    v.surfaces = s_surf;
    v.fast = s_fast;
    v.sv = A.sv;
    v.tables32 = A.tables32;
    FRay r = load_fray(in, i);
    // a pass-through surface 0 is elided as in the grid kernels
    const bool skip = s_spec->skip_first != 0;
    const int k0 = skip ? 1 : 0;
    if (skip) {
      const bool parallel = r.dz == 0.0;
      fast_xf(s_spec->first_skip, r);
      if (parallel) {
        r.px = r.py = r.pz = NAN;
        r.intensity = NAN;
      }
    } else {
      fast_xf(s_spec->first_hop[1], r);
    }
    for (int k = k0; k < A.num_surfaces; ++k) {
      if (k > k0) {
        fast_xf(s_fast[k].hop, r);
      }
      fast_propagate<F>(s_fast[k], s_surf[k], v, path, r);
    }
    const optk_affine &last = s_spec->last[A.apply_post];
    fast_xf(last, r);
    store_f(out, (int64_t)path * A.n + i, r);
    // statistics of the float32 result, as a caller reading the outputs sees it
    x = (double)(float)r.px;
    y = (double)(float)r.py;
    inten = r.intensity;
    unv = r.unvignetted;
  }
};

// the same with the instances' surface kinds fixed at compile time (every
// instance of a batch matching one signature), the surfaces unrolled
template <int F, uint32_t K>
__device__ __forceinline__ void batch_fast_step(int k, int k0, const FastSurface *s_fast,
                                                const optk_surface *s_surf, const FastView &v, int path,
                                                FRay &r) {
  if (k < k0) return;
  if (k > k0) fast_xf(s_fast[k].hop, r);
  fast_propagate<F, K>(s_fast[k], s_surf[k], v, path, r);
}

template <int F, uint32_t... K> struct Tracer32Spec : Tracer32<F> {
  __device__ __forceinline__ void trace(int64_t i, int path, double &x, double &y, double &inten,
                                        bool &unv) {
    FastView v;
    v.surfaces = this->s_surf;
    v.fast = this->s_fast;
    v.sv = this->A.sv;
    v.tables32 = this->A.tables32;
    FRay r = load_fray(this->in, i);
    const bool skip = this->s_spec->skip_first != 0;
    const int k0 = skip ? 1 : 0;
    if (skip) {
      const bool parallel = r.dz == 0.0;
      fast_xf(this->s_spec->first_skip, r);
      if (parallel) {
        r.px = r.py = r.pz = NAN;
        r.intensity = NAN;
      }
    } else {
      fast_xf(this->s_spec->first_hop[1], r);
    }
    int k = 0;
    (batch_fast_step<F, K>(k++, k0, this->s_fast, this->s_surf, v, path, r), ...);
    const optk_affine &last = this->s_spec->last[this->A.apply_post];
    fast_xf(last, r);
    store_f(this->out, (int64_t)path * this->A.n + i, r);
    x = (double)(float)r.px;
    y = (double)(float)r.py;
    inten = r.intensity;
    unv = r.unvignetted;
  }
};

template <int F>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(BatchWaves<F>::value)))
void batch_grid_f32(const BatchArgs A, const FGridIn in, const FRaysOut out) {
  extern __shared__ double staged[];
  FastSpec *spec = reinterpret_cast<FastSpec *>(staged);
  FastSurface *fast = reinterpret_cast<FastSurface *>(spec + 1);
  optk_surface *surf = reinterpret_cast<optk_surface *>(fast + A.num_surfaces);
  Tracer32<F> tr{A, in, out, spec, fast, surf};
  batch_loop(A, tr);
}

template <int F, uint32_t... K>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(FastSpecWaves<F>::value)))
void batch_grid_f32_spec(const BatchArgs A, const FGridIn in, const FRaysOut out) {
  extern __shared__ double staged[];
  FastSpec *spec = reinterpret_cast<FastSpec *>(staged);
  FastSurface *fast = reinterpret_cast<FastSurface *>(spec + 1);
  optk_surface *surf = reinterpret_cast<optk_surface *>(fast + A.num_surfaces);
  Tracer32Spec<F, K...> tr{{A, in, out, spec, fast, surf}};
  batch_loop(A, tr);
}

typedef void (*BatchSpec32)(BatchArgs, FGridIn, FRaysOut);
// by signature; the geometric Newtonian layout keeps the generic kernel
const BatchSpec32 BATCH_SPEC32[NUM_SIGNATURES] = {
    nullptr,
    batch_grid_f32_spec<SIGNATURES[SIG_SPECTROGRAPH].features, OPTK_SIGNATURE_SPECTROGRAPH>,
    batch_grid_f32_spec<SIGNATURES[SIG_SPECTROGRAPH_TILTED].features, OPTK_SIGNATURE_SPECTROGRAPH_TILTED>,
};

void *batch_kernel_f32(int features, int signature) {
  if (signature >= 0 && signature < NUM_SIGNATURES && BATCH_SPEC32[signature])
    return (void *)BATCH_SPEC32[signature];
  if (features == 0) return (void *)batch_grid_f32<0>;
  if (features == (FEAT_TOROID | FEAT_RULINGS)) return (void *)batch_grid_f32<FEAT_TOROID | FEAT_RULINGS>;
  return (void *)batch_grid_f32<FEAT_ALL>;
}

}  // namespace

namespace optk {

int32_t fast_trace_rays(const optk_system *sys, const FRaysIn &in, const FRaysOut &out,
                        const FRaysOut *accum, int64_t n, int32_t apply_post, bool generic,
                        hipStream_t stream) {
  static const KernelSet<FRaysIn> ks = {
      {trace_rays_f32<0, false>, trace_rays_f32<FEAT_TOROID | FEAT_RULINGS, false>,
       trace_rays_f32<FEAT_ALL, false>, trace_rays_f32<FEAT_ALL, true>},
      {trace_rays_f32<0, false, IO_FULL>, trace_rays_f32<FEAT_TOROID | FEAT_RULINGS, false, IO_FULL>,
       trace_rays_f32<FEAT_ALL, false, IO_FULL>}};
  const SpecRays32 spec = (sys->signature >= 0 && !generic) ? SPEC_RAYS32[sys->signature] : nullptr;
  return launch<FRaysIn>(ks, full_output(out), sys, in, out, accum, n, 0, apply_post, spec, stream);
}

int32_t fast_trace_grid(const optk_system *sys, const FGridIn &in, const FRaysOut &out,
                        const FRaysOut *accum, int64_t n, int32_t apply_post, bool generic,
                        hipStream_t stream) {
  static const KernelSet<FGridIn> ks = {
      {trace_grid_f32<0, false>, trace_grid_f32<FEAT_TOROID | FEAT_RULINGS, false>,
       trace_grid_f32<FEAT_ALL, false>, trace_grid_f32<FEAT_ALL, true>},
      {trace_grid_f32<0, false, IO_SENSOR>, trace_grid_f32<FEAT_TOROID | FEAT_RULINGS, false, IO_SENSOR>,
       trace_grid_f32<FEAT_ALL, false, IO_SENSOR>}};
  const SpecGrid32 spec = (sys->signature >= 0 && !generic) ? SPEC_GRID32[sys->signature] : nullptr;
  return launch<FGridIn>(ks, sensor_layout(in, out), sys, in, out, accum, n, 1, apply_post, spec,
                         stream);
}

size_t batch_lds_f32(int num_surfaces) { return fast_staged_bytes(num_surfaces); }

int batch_occupancy_f32(int features, int signature, int block, size_t lds) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_kernel_f32(features, signature), block,
                                                   lds) != hipSuccess)
    return 0;
  return nb;
}

hipError_t batch_launch_f32(int features, int signature, int blocks, int block, const BatchArgs &A,
                            const FGridIn &in, const FRaysOut &out, hipStream_t stream) {
  size_t lds = fast_staged_bytes(A.num_surfaces);
  if (signature >= 0 && signature < NUM_SIGNATURES && BATCH_SPEC32[signature])
    hipLaunchKernelGGL(BATCH_SPEC32[signature], dim3(blocks), dim3(block), lds, stream, A, in, out);
  else if (features == 0)
    hipLaunchKernelGGL(batch_grid_f32<0>, dim3(blocks), dim3(block), lds, stream, A, in, out);
  else if (features == (FEAT_TOROID | FEAT_RULINGS))
    hipLaunchKernelGGL(batch_grid_f32<FEAT_TOROID | FEAT_RULINGS>, dim3(blocks), dim3(block), lds,
                       stream, A, in, out);
  else
    hipLaunchKernelGGL(batch_grid_f32<FEAT_ALL>, dim3(blocks), dim3(block), lds, stream, A, in, out);
  return hipGetLastError();
}

}  // namespace optk
