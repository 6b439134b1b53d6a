// optk_batch.hip -- host side of the batched persistent trace (optk_batch_*):
// per-instance records, the work-tile table, the launch-shape autotuner and
// hipGraph capture/replay.  The persistent kernels themselves live with
// their forms (optk_lib.hip: float64, optk_fast.hip: float32).

#include <cstring>
#include <string>
#include <vector>

#include "optk_internal.h"
#include "optk_signatures.h"

using namespace optk;

namespace {

__global__ __launch_bounds__(256) void batch_zero(double *stats, int64_t count, int32_t *counter) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *counter = 0;
  for (; i < count; i += (int64_t)gridDim.x * blockDim.x) stats[i] = 0.0;
}

// RMS spot radius about the centroid and throughput per (instance, path)
__global__ __launch_bounds__(256) void batch_finalize(double *stats, const int64_t *counts,
                                                      int32_t num_instances, int32_t num_paths) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (int64_t)num_instances * num_paths) return;
  double *s = stats + k * OPTK_BATCH_STATS;
  const double n = s[0];
  const double mx = s[2] / n, my = s[3] / n;
  const double var = (s[4] / n - mx * mx) + (s[5] / n - my * my);
  s[6] = sqrt(var > 0.0 ? var : 0.0);
  if (n == 0.0) s[6] = NAN;
  const int64_t rays = counts[k / num_paths];
  s[7] = rays > 0 ? s[1] / (double)rays : 0.0;
}

struct LaunchShape {
  int block, rays_per_thread, blocks;
};

}  // namespace

struct optk_batch {
  int32_t num_instances = 0;
  int32_t num_surfaces = 0;
  int32_t num_paths = 1;
  int32_t features = 0;
  int32_t signature = -1;  // the instances' common surface-kind signature, or -1
  int32_t num_cus = 0;
  // shared by every instance
  optk_layer *layers = nullptr;
  optk_stack *stacks = nullptr;
  optk_chemical *chemicals = nullptr;
  double *tables = nullptr;
  float *tables32 = nullptr;
  ChemGrid *grids = nullptr;
  int32_t *grid_lo = nullptr;
  // per instance
  optk_surface *surfaces = nullptr;
  TraceSpec *specs64 = nullptr;
  FastSurface *fast = nullptr;
  FastSpec *fast_specs = nullptr;
  // work queue and tiles
  int32_t *counter = nullptr;
  int64_t *counts = nullptr;  // rays per instance (device)
  BatchTile *tiles = nullptr;
  int64_t tiles_capacity = 0;
  int64_t num_tiles = 0;
  int64_t tile_rays = 0;
  std::vector<int64_t> offsets;  // the offsets the tile table was built for
  // graph replay
  hipStream_t capture = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<uint8_t> exec_key;
  // autotuned shapes per precision
  optk_tune tuned[2];
  bool have_tuned[2] = {false, false};
};

namespace {

template <typename T> int32_t upload(T **dst, const T *src, int64_t count) {
  *dst = nullptr;
  if (count <= 0) return OPTK_OK;
  hipError_t e = hipMalloc((void **)dst, sizeof(T) * count);
  if (e != hipSuccess) return fail(OPTK_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  e = hipMemcpy(*dst, src, sizeof(T) * count, hipMemcpyHostToDevice);
  if (e != hipSuccess) return fail(OPTK_EDEVICE, std::string("hipMemcpy: ") + hipGetErrorString(e));
  return OPTK_OK;
}

void release(optk_batch *b) {
  (void)hipFree(b->layers);
  (void)hipFree(b->stacks);
  (void)hipFree(b->chemicals);
  (void)hipFree(b->tables);
  (void)hipFree(b->tables32);
  (void)hipFree(b->grids);
  (void)hipFree(b->grid_lo);
  (void)hipFree(b->surfaces);
  (void)hipFree(b->specs64);
  (void)hipFree(b->fast);
  (void)hipFree(b->fast_specs);
  (void)hipFree(b->counter);
  (void)hipFree(b->counts);
  (void)hipFree(b->tiles);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->capture) (void)hipStreamDestroy(b->capture);
}

// (re)build the tile table for these offsets and this tile size
int32_t prepare_tiles(optk_batch *b, const int64_t *offsets, int64_t tile_rays) {
  const int n = b->num_instances;
  bool same = b->tile_rays == tile_rays && (int64_t)b->offsets.size() == n + 1 &&
              std::memcmp(b->offsets.data(), offsets, sizeof(int64_t) * (n + 1)) == 0;
  if (same) return OPTK_OK;
  std::vector<BatchTile> tiles;
  std::vector<int64_t> counts(n);
  for (int i = 0; i < n; ++i) {
    const int64_t lo = offsets[i], hi = offsets[i + 1];
    counts[i] = hi - lo;
    for (int64_t s = lo; s < hi; s += tile_rays) {
      BatchTile t;
      t.start = s;
      t.instance = i;
      t.count = (int32_t)((hi - s) < tile_rays ? (hi - s) : tile_rays);
      tiles.push_back(t);
    }
  }
  if (tiles.size() > (size_t)INT32_MAX) return fail(OPTK_EINVAL, "too many work tiles");
  if ((int64_t)tiles.size() > b->tiles_capacity) {
    (void)hipFree(b->tiles);
    b->tiles = nullptr;
    b->tiles_capacity = 0;
    hipError_t e = hipMalloc((void **)&b->tiles, sizeof(BatchTile) * (tiles.size() + 1));
    if (e != hipSuccess) return fail(OPTK_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    b->tiles_capacity = (int64_t)tiles.size() + 1;
  }
  // blocking copies: the table changes only with new offsets or tile sizes
  if (!tiles.empty()) {
    hipError_t e = hipMemcpy(b->tiles, tiles.data(), sizeof(BatchTile) * tiles.size(),
                             hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(OPTK_EDEVICE, std::string("hipMemcpy: ") + hipGetErrorString(e));
  }
  hipError_t e = hipMemcpy(b->counts, counts.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice);
  if (e != hipSuccess) return fail(OPTK_EDEVICE, std::string("hipMemcpy: ") + hipGetErrorString(e));
  b->num_tiles = (int64_t)tiles.size();
  b->tile_rays = tile_rays;
  b->offsets.assign(offsets, offsets + n + 1);
  // a graph captured for the old table would replay it
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
    b->exec_key.clear();
  }
  return OPTK_OK;
}

int32_t check_tune(const optk_tune &t) {
  if (t.block != 64 && t.block != 128 && t.block != 256 && t.block != 512)
    return fail(OPTK_EINVAL, "tune.block must be 64, 128, 256 or 512");
  if (t.rays_per_thread < 1 || t.rays_per_thread > 64)
    return fail(OPTK_EINVAL, "tune.rays_per_thread must be in 1..64");
  if (t.blocks_per_cu < 0) return fail(OPTK_EINVAL, "tune.blocks_per_cu must be >= 0");
  return OPTK_OK;
}

struct Inputs {
  GridIn g64;
  RaysOut o64;
  FGridIn g32;
  FRaysOut o32;
};

void fill_inputs(Inputs &io, int32_t precision, const optk_grid *in, const optk_rays *out) {
  std::memset(&io, 0, sizeof(io));
  for (int k = 0; k < 2; ++k) {
    io.g64.field[k] = (const double *)in->field[k];
    io.g64.pupil[k] = (const double *)in->pupil[k];
    io.g64.field_default[k] = in->field_default[k];
    io.g64.pupil_default[k] = in->pupil_default[k];
    io.g32.field[k] = (const float *)in->field[k];
    io.g32.pupil[k] = (const float *)in->pupil[k];
    io.g32.field_default[k] = (float)in->field_default[k];
    io.g32.pupil_default[k] = (float)in->pupil_default[k];
  }
  io.g64.wavelength = (const double *)in->wavelength;
  io.g64.intensity = (const double *)in->intensity;
  io.g64.wavelength_default = in->wavelength_default;
  io.g64.intensity_default = in->intensity_default;
  io.g64.object_at_infinity = in->object_at_infinity;
  io.g32.wavelength = (const float *)in->wavelength;
  io.g32.intensity = (const float *)in->intensity;
  io.g32.wavelength_default = (float)in->wavelength_default;
  io.g32.intensity_default = (float)in->intensity_default;
  io.g32.object_at_infinity = in->object_at_infinity;
  if (out) {
    for (int k = 0; k < 3; ++k) {
      io.o64.p[k] = (double *)out->position[k];
      io.o64.d[k] = (double *)out->direction[k];
      io.o32.p[k] = (float *)out->position[k];
      io.o32.d[k] = (float *)out->direction[k];
    }
    io.o64.wavelength = (double *)out->wavelength;
    io.o64.intensity = (double *)out->intensity;
    io.o64.attenuation = (double *)out->attenuation;
    io.o64.index = (double *)out->index_refraction;
    io.o64.unvignetted = out->unvignetted;
    io.o32.wavelength = (float *)out->wavelength;
    io.o32.intensity = (float *)out->intensity;
    io.o32.attenuation = (float *)out->attenuation;
    io.o32.index = (float *)out->index_refraction;
    io.o32.unvignetted = out->unvignetted;
  }
  (void)precision;
}

// the signature whose specialised kernels a call runs (-1: the generic ones)
int call_signature(const optk_batch *b, int32_t flags) {
  return (flags & OPTK_TRACE_GENERIC) ? -1 : b->signature;
}

int resident_blocks(const optk_batch *b, int32_t precision, int block, int32_t flags) {
  const int sig = call_signature(b, flags);
  if (precision == OPTK_F64)
    return batch_occupancy_f64(b->features, sig, block, batch_lds_f64(b->num_surfaces));
  return batch_occupancy_f32(b->features, sig, block, batch_lds_f32(b->num_surfaces));
}

// zero the statistics and the queue head, run the persistent kernel, reduce
// the statistics: three launches on `stream`, no host synchronisation
int32_t enqueue(optk_batch *b, int32_t precision, const Inputs &io, double *stats,
                const LaunchShape &shape, int32_t flags, int64_t n, hipStream_t stream) {
  const int64_t nstats = (int64_t)b->num_instances * b->num_paths * OPTK_BATCH_STATS;
  hipLaunchKernelGGL(batch_zero, dim3(64), dim3(256), 0, stream, stats, nstats, b->counter);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(OPTK_EDEVICE, std::string("batch zero: ") + hipGetErrorString(e));
  BatchArgs A;
  std::memset(&A, 0, sizeof(A));
  A.tiles = b->tiles;
  A.num_tiles = b->num_tiles;
  A.counter = b->counter;
  A.rays_per_thread = shape.rays_per_thread;
  A.num_surfaces = b->num_surfaces;
  A.num_paths = b->num_paths;
  A.apply_post = (flags & OPTK_TRACE_APPLY_POST) ? 1 : 0;
  A.surfaces = b->surfaces;
  A.specs64 = b->specs64;
  A.fast = b->fast;
  A.fast_specs = b->fast_specs;
  A.sv.grids = b->grids;
  A.sv.grid_lo = b->grid_lo;
  A.sv.layers = b->layers;
  A.sv.stacks = b->stacks;
  A.sv.chemicals = b->chemicals;
  A.sv.tables = b->tables;
  A.tables32 = b->tables32;
  A.stats = stats;
  A.n = n;
  if (b->num_tiles > 0) {
    const int sig = call_signature(b, flags);
    e = precision == OPTK_F64
            ? batch_launch_f64(b->features, sig, shape.blocks, shape.block, A, io.g64, io.o64, stream)
            : batch_launch_f32(b->features, sig, shape.blocks, shape.block, A, io.g32, io.o32, stream);
    if (e != hipSuccess) return fail(OPTK_EDEVICE, std::string("batch trace: ") + hipGetErrorString(e));
  }
  const int64_t pairs = (int64_t)b->num_instances * b->num_paths;
  hipLaunchKernelGGL(batch_finalize, dim3((unsigned)((pairs + 255) / 256)), dim3(256), 0, stream,
                     stats, b->counts, b->num_instances, b->num_paths);
  e = hipGetLastError();
  if (e != hipSuccess) return fail(OPTK_EDEVICE, std::string("batch finalize: ") + hipGetErrorString(e));
  return OPTK_OK;
}

int32_t check_call(const optk_batch *b, int32_t precision, const optk_grid *in,
                   const int64_t *offsets, double *stats) {
  if (!b) return fail(OPTK_EINVAL, "null batch");
  if (precision != OPTK_F64 && precision != OPTK_F32) return fail(OPTK_EINVAL, "bad precision");
  if (!in || !offsets || !stats) return fail(OPTK_EINVAL, "null argument");
  if (offsets[0] != 0) return fail(OPTK_EINVAL, "offsets[0] must be 0");
  for (int i = 0; i < b->num_instances; ++i)
    if (offsets[i + 1] < offsets[i]) return fail(OPTK_EINVAL, "offsets must be non-decreasing");
  return OPTK_OK;
}

int32_t shape_for(optk_batch *b, int32_t precision, const optk_tune &t, int32_t flags,
                  LaunchShape &shape) {
  int32_t rc = check_tune(t);
  if (rc) return rc;
  int fit = resident_blocks(b, precision, t.block, flags);
  if (fit <= 0) return fail(OPTK_EDEVICE, "the batched kernel does not fit on a CU at this block size");
  int per_cu = (t.blocks_per_cu > 0 && t.blocks_per_cu < fit) ? t.blocks_per_cu : fit;
  shape.block = t.block;
  shape.rays_per_thread = t.rays_per_thread;
  shape.blocks = per_cu * b->num_cus;
  return OPTK_OK;
}

}  // namespace

extern "C" {

int32_t optk_batch_destroy(optk_batch *batch) {
  if (!batch) return OPTK_OK;
  release(batch);
  delete batch;
  return OPTK_OK;
}

int32_t optk_batch_create(const optk_batch_desc *desc, optk_batch **out) {
  if (!desc || !out || !desc->base) return fail(OPTK_EINVAL, "null argument");
  *out = nullptr;
  const optk_system_desc &base = *desc->base;
  const int N = desc->num_instances;
  const int S = base.num_surfaces;
  if (N < 1) return fail(OPTK_EINVAL, "num_instances must be >= 1");
  if (S > 0 && !desc->surfaces) return fail(OPTK_EINVAL, "null surface records");
  if ((base.num_layers > 0 && !base.layers) || (base.num_stacks > 0 && !base.stacks) ||
      (base.num_chemicals > 0 && !base.chemicals) || (base.num_tables > 0 && !base.tables))
    return fail(OPTK_EINVAL, "null record array");
  // every instance is a valid system over the shared tables; all instances
  // need the same kernel variant
  std::vector<optk_surface> surfaces(desc->surfaces, desc->surfaces + (size_t)N * S);
  int32_t features = -1;
  for (int i = 0; i < N; ++i) {
    optk_system_desc d = base;
    d.surfaces = surfaces.data() + (size_t)i * S;
    if (desc->pre) {
      d.pre[0] = desc->pre[2 * i];
      d.pre[1] = desc->pre[2 * i + 1];
    }
    if (desc->post) d.post = desc->post[i];
    int32_t rc = validate_desc(&d);
    if (rc) return fail(rc, "instance " + std::to_string(i) + ": " + optk_last_error());
    int32_t f = system_features(&d);
    if (features >= 0 && f != features)
      return fail(OPTK_EINVAL, "instances need different kernel variants (surface kinds differ)");
    features = f;
  }
  for (auto &s : surfaces) derive_constants(s);
  // the specialised kernels serve a batch whose instances all match one
  // signature (perturbations keep the kinds of every record)
  int32_t signature = find_signature(surfaces.data(), S, features);
  for (int i = 1; i < N && signature >= 0; ++i)
    if (find_signature(surfaces.data() + (size_t)i * S, S, features) != signature) signature = -1;

  optk_batch *b = new optk_batch();
  b->num_instances = N;
  b->num_surfaces = S;
  b->num_paths = base.num_paths;
  b->features = features;
  b->signature = signature;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&b->num_cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) e = hipStreamCreate(&b->capture);
  if (e != hipSuccess) {
    release(b);
    delete b;
    return fail(OPTK_EDEVICE, std::string("device setup: ") + hipGetErrorString(e));
  }
  std::vector<float> tables32(base.tables, base.tables + base.num_tables);
  int32_t rc;
  if ((rc = upload(&b->layers, base.layers, base.num_layers)) ||
      (rc = upload(&b->stacks, base.stacks, base.num_stacks)) ||
      (rc = upload(&b->chemicals, base.chemicals, base.num_chemicals)) ||
      (rc = upload(&b->tables, base.tables, base.num_tables)) ||
      (rc = upload(&b->tables32, tables32.data(), base.num_tables)) ||
      (rc = upload_chem_grids(base.chemicals, base.num_chemicals, base.tables, &b->grids,
                              &b->grid_lo)) ||
      (rc = upload(&b->surfaces, surfaces.data(), (int64_t)N * S))) {
    release(b);
    delete b;
    return rc;
  }
  std::vector<TraceSpec> specs(N);
  std::vector<FastSurface> fast((size_t)N * (S > 0 ? S : 1));
  std::vector<FastSpec> fast_specs(N);
  for (int i = 0; i < N; ++i) {
    TraceSpec &sp = specs[i];
    std::memset(&sp, 0, sizeof(sp));
    sp.sys.surfaces = nullptr;  // the kernel points it at the staged copy
    sp.sys.num_surfaces = S;
    sp.sys.num_paths = base.num_paths;
    sp.sys.sv.grids = b->grids;
    sp.sys.sv.grid_lo = b->grid_lo;
    sp.sys.sv.layers = b->layers;
    sp.sys.sv.stacks = b->stacks;
    sp.sys.sv.chemicals = b->chemicals;
    sp.sys.sv.tables = b->tables;
    sp.pre[0] = desc->pre ? desc->pre[2 * i] : base.pre[0];
    sp.pre[1] = desc->pre ? desc->pre[2 * i + 1] : base.pre[1];
    sp.pre_is_inverse[0] = base.pre_is_inverse[0];
    sp.pre_is_inverse[1] = base.pre_is_inverse[1];
    sp.post = desc->post ? desc->post[i] : base.post;
    std::memset(&fast_specs[i], 0, sizeof(FastSpec));
    fast_build(surfaces.data() + (size_t)i * S, S, sp.pre, sp.pre_is_inverse, sp.post,
               fast.data() + (size_t)i * S, fast_specs[i]);
  }
  if ((rc = upload(&b->specs64, specs.data(), N)) ||
      (rc = upload(&b->fast, fast.data(), (int64_t)N * S)) ||
      (rc = upload(&b->fast_specs, fast_specs.data(), N))) {
    release(b);
    delete b;
    return rc;
  }
  e = hipMalloc((void **)&b->counter, sizeof(int32_t));
  if (e == hipSuccess) e = hipMalloc((void **)&b->counts, sizeof(int64_t) * N);
  if (e != hipSuccess) {
    release(b);
    delete b;
    return fail(OPTK_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  *out = b;
  return OPTK_OK;
}

int32_t optk_batch_trace(optk_batch *b, int32_t precision, const optk_grid *in,
                         const int64_t *offsets, const optk_rays *out, double *stats,
                         const optk_tune *tune, int32_t flags, void *hip_stream) {
  int32_t rc = check_call(b, precision, in, offsets, stats);
  if (rc) return rc;
  hipStream_t stream = (hipStream_t)hip_stream;
  optk_tune t;
  if (tune)
    t = *tune;
  else if (b->have_tuned[precision])
    t = b->tuned[precision];
  else
    t = optk_tune{256, 2, 0, 0, 0.0};
  LaunchShape shape;
  if ((rc = shape_for(b, precision, t, flags, shape))) return rc;
  if ((rc = prepare_tiles(b, offsets, (int64_t)shape.block * shape.rays_per_thread))) return rc;
  const int64_t n = offsets[b->num_instances];
  Inputs io;
  fill_inputs(io, precision, in, out);
  if (!(flags & OPTK_BATCH_GRAPH)) return enqueue(b, precision, io, stats, shape, flags, n, stream);

  // graph: replay while the launch is identical, else capture it again
  std::vector<uint8_t> key(sizeof(io) + sizeof(shape) + 3 * sizeof(int64_t) + sizeof(void *));
  uint8_t *k = key.data();
  std::memcpy(k, &io, sizeof(io));
  k += sizeof(io);
  std::memcpy(k, &shape, sizeof(shape));
  k += sizeof(shape);
  int64_t extra[3] = {precision, flags, n};
  std::memcpy(k, extra, sizeof(extra));
  k += sizeof(extra);
  std::memcpy(k, &stats, sizeof(void *));
  if (!b->exec || b->exec_key != key) {
    if (b->exec) {
      (void)hipGraphExecDestroy(b->exec);
      b->exec = nullptr;
    }
    hipError_t e = hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return fail(OPTK_EDEVICE, std::string("begin capture: ") + hipGetErrorString(e));
    rc = enqueue(b, precision, io, stats, shape, flags, n, b->capture);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(b->capture, &graph);
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (e != hipSuccess) return fail(OPTK_EDEVICE, std::string("end capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
      b->exec = nullptr;
      return fail(OPTK_EDEVICE, std::string("graph instantiate: ") + hipGetErrorString(e));
    }
    b->exec_key = key;
  }
  hipError_t e = hipGraphLaunch(b->exec, stream);
  if (e != hipSuccess) return fail(OPTK_EDEVICE, std::string("graph launch: ") + hipGetErrorString(e));
  return OPTK_OK;
}

int32_t optk_batch_autotune(optk_batch *b, int32_t precision, const optk_grid *in,
                            const int64_t *offsets, double *stats, int32_t flags,
                            void *hip_stream, optk_tune *best) {
  int32_t rc = check_call(b, precision, in, offsets, stats);
  if (rc) return rc;
  if (!best) return fail(OPTK_EINVAL, "null result");
  hipStream_t stream = (hipStream_t)hip_stream;
  const int64_t n = offsets[b->num_instances];
  Inputs io;
  fill_inputs(io, precision, in, nullptr);  // timing runs write statistics only
  hipEvent_t t0, t1;
  hipError_t e = hipEventCreate(&t0);
  if (e == hipSuccess) e = hipEventCreate(&t1);
  if (e != hipSuccess) return fail(OPTK_EDEVICE, std::string("events: ") + hipGetErrorString(e));
  static const int blocks[] = {64, 128, 256, 512};
  static const int rpts[] = {1, 2, 4, 8};
  optk_tune found = {0, 0, 0, 0, INFINITY};
  for (int bi = 0; bi < 4 && rc == OPTK_OK; ++bi) {
    for (int ri = 0; ri < 4 && rc == OPTK_OK; ++ri) {
      optk_tune t = {blocks[bi], rpts[ri], 0, 0, 0.0};
      LaunchShape shape;
      if (shape_for(b, precision, t, flags, shape) != OPTK_OK) continue;  // does not fit: skip
      if ((rc = prepare_tiles(b, offsets, (int64_t)shape.block * shape.rays_per_thread))) break;
      if ((rc = enqueue(b, precision, io, stats, shape, flags, n, stream))) break;  // warm-up
      float best_ms = INFINITY;
      for (int rep = 0; rep < 2 && rc == OPTK_OK; ++rep) {
        e = hipEventRecord(t0, stream);
        if (e == hipSuccess) rc = enqueue(b, precision, io, stats, shape, flags, n, stream);
        if (e == hipSuccess && rc == OPTK_OK) e = hipEventRecord(t1, stream);
        if (e == hipSuccess && rc == OPTK_OK) e = hipEventSynchronize(t1);
        float ms = 0.0f;
        if (e == hipSuccess && rc == OPTK_OK) e = hipEventElapsedTime(&ms, t0, t1);
        if (e != hipSuccess && rc == OPTK_OK)
          rc = fail(OPTK_EDEVICE, std::string("autotune timing: ") + hipGetErrorString(e));
        if (ms < best_ms) best_ms = ms;
      }
      if (rc == OPTK_OK && best_ms < found.ms) {
        found = t;
        found.ms = best_ms;
      }
    }
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if (rc) return rc;
  if (found.block == 0) return fail(OPTK_EDEVICE, "no launch shape fits");
  b->tuned[precision] = found;
  b->have_tuned[precision] = true;
  *best = found;
  return OPTK_OK;
}

}  // extern "C"
