// optk_internal.h -- state shared by the translation units of liboptk.so
// (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "optk_fast.h"

// A compiled system: the optk_surface records and their tables on the
// device, plus the derived records of the float32 form.
struct optk_system {
  optk_surface *surfaces = nullptr;
  optk_layer *layers = nullptr;
  optk_stack *stacks = nullptr;
  optk_chemical *chemicals = nullptr;
  double *tables = nullptr;
  int32_t num_surfaces = 0;
  int32_t num_paths = 1;
  int32_t num_layers = 0;
  int32_t num_stacks = 0;
  int32_t num_chemicals = 0;
  int64_t num_tables = 0;
  optk_affine pre[2];
  int32_t pre_is_inverse[2];
  optk_affine post;
  int32_t features = 0;                  // FEAT_* the surfaces need
  int32_t signature = -1;                // specialised kernels (optk_signatures.h), or -1
  optk::TraceSpec *spec = nullptr;       // device copy of views + transformations
  std::vector<optk_surface> host_surfaces;
  // float32 form
  optk::FastSurface *fast = nullptr;     // one per surface
  optk::FastSpec *fast_spec = nullptr;
  float *tables32 = nullptr;             // tables[] rounded to float32
  optk::ChemGrid *grids = nullptr;       // n/k lookup grids, one per chemical
  int32_t *grid_lo = nullptr;
};

namespace optk {

// record the thread's last error message and return `code`
int32_t fail(int32_t code, const std::string &msg);

// Blocks of a grid-stride launch of `kernel` (block threads, no dynamic
// LDS) over n items: four full rounds of the blocks the current device
// holds resident at once (CUs x blocks per CU, cached per kernel and
// device) -- whole rounds, so no partly filled last round idles CUs, and
// more than one, so blocks that finish early are replaced (measured on the
// geometric traces: 1 round 4-6 % slower than 4, 8 rounds no faster);
// fewer when n needs fewer.
int64_t resident_grid(const void *kernel, int block, int64_t n);

// threads per block of the trace kernels (grid and ray entry points, both
// forms); their grid-stride loops use it as a compile-time constant
constexpr int TRACE_BLOCK = 256;

// check every index and range a kernel will follow (OPTK_OK or an error)
int32_t validate_desc(const optk_system_desc *d);
// the kernel feature variant (FEAT_*) covering the system's surfaces
int32_t system_features(const optk_system_desc *d);

// upload the n/k lookup grids of the chemicals (host tables) to the device
int32_t upload_chem_grids(const optk_chemical *chems, int num_chemicals, const double *tables,
                          ChemGrid **grids, int32_t **grid_lo);

// float64 ray buffers (device pointers; null = use the default)
struct RaysIn {
  const double *p[3];
  const double *d[3];
  const double *wavelength, *intensity, *attenuation, *index;
  const uint8_t *unvignetted;
  double wavelength_default, intensity_default, attenuation_default, index_default;
  int32_t unvignetted_default;
};

struct RaysOut {
  double *p[3];
  double *d[3];
  double *wavelength, *intensity, *attenuation, *index;
  uint8_t *unvignetted;
};

struct GridIn {
  const double *field[2];
  const double *pupil[2];
  const double *wavelength, *intensity;
  double field_default[2], pupil_default[2], wavelength_default, intensity_default;
  int32_t object_at_infinity;
};

// float32 ray buffers (device pointers; null = use the default)
struct FRaysIn {
  const float *p[3];
  const float *d[3];
  const float *wavelength, *intensity, *attenuation, *index;
  const uint8_t *unvignetted;
  float wavelength_default, intensity_default, attenuation_default, index_default;
  int32_t unvignetted_default;
};

struct FRaysOut {
  float *p[3];
  float *d[3];
  float *wavelength, *intensity, *attenuation, *index;
  uint8_t *unvignetted;
};

struct FGridIn {
  const float *field[2];
  const float *pupil[2];
  const float *wavelength, *intensity;
  float field_default[2], pupil_default[2], wavelength_default, intensity_default;
  int32_t object_at_infinity;
};

// float32-form launches (optk_fast.hip); `accum` may be null
// `generic`: the generic kernels even where the system has specialised ones
int32_t fast_trace_rays(const optk_system *sys, const FRaysIn &in, const FRaysOut &out,
                        const FRaysOut *accum, int64_t n, int32_t apply_post, bool generic,
                        hipStream_t stream);
int32_t fast_trace_grid(const optk_system *sys, const FGridIn &in, const FRaysOut &out,
                        const FRaysOut *accum, int64_t n, int32_t apply_post, bool generic,
                        hipStream_t stream);

// ---------------------------------------------------------------------------
// batched persistent trace (optk_batch_*)
// ---------------------------------------------------------------------------

// a work tile: `count` consecutive rays of one instance
struct BatchTile {
  int64_t start;
  int32_t instance;
  int32_t count;
};

struct BatchArgs {
  const BatchTile *tiles;
  int64_t num_tiles;
  int32_t *counter;                // work-queue head, zeroed before each launch
  int32_t rays_per_thread;
  int32_t num_surfaces;
  int32_t num_paths;
  int32_t apply_post;
  const optk_surface *surfaces;    // [instance][surface]
  const TraceSpec *specs64;        // [instance]
  const FastSurface *fast;         // [instance][surface]
  const FastSpec *fast_specs;      // [instance]
  StackView sv;
  const float *tables32;
  double *stats;                   // [instance][path][OPTK_BATCH_STATS]
  int64_t n;                       // rays in the batch (output stride per path)
};

size_t batch_lds_f64(int num_surfaces);
size_t batch_lds_f32(int num_surfaces);
// resident blocks per CU of the kernel covering `features` (0 on error);
// signature >= 0: the instances' common surface-kind signature
// (optk_signatures.h), whose specialised kernel runs where one exists
int batch_occupancy_f64(int features, int signature, int block, size_t lds);
int batch_occupancy_f32(int features, int signature, int block, size_t lds);
hipError_t batch_launch_f64(int features, int signature, int blocks, int block, const BatchArgs &A,
                            const GridIn &in, const RaysOut &out, hipStream_t stream);
hipError_t batch_launch_f32(int features, int signature, int blocks, int block, const BatchArgs &A,
                            const FGridIn &in, const FRaysOut &out, hipStream_t stream);

}  // namespace optk
