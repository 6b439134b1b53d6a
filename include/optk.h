/*
 * optk.h -- C ABI of the MI355X sequential ray tracer (liboptk.so).
 *
 * Plain pointers and sizes only.  All ray buffers are caller-owned device
 * memory laid out structure-of-arrays (one contiguous array per component);
 * the library owns only compiled system descriptors.  Every entry point
 * returns 0 on success or a negative OPTK_E* code, never throws, and leaves a
 * message for optk_last_error() (thread-local).  Work is enqueued on the
 * caller's HIP stream; no entry point synchronises the device except
 * optk_system_create/destroy (host<->device copies of the descriptor).
 *
 * Units: lengths mm, wavelengths nm, angles in radians inside records (the
 * host layer converts degrees), attenuation 1/mm.
 *
 * Reference interface each entry point replaces (sun-data/optika):
 *   optk_trace       optika/propagators.py:19-74 (propagate_rays /
 *                    accumulate_rays over AbstractSurface.propagate_rays,
 *                    optika/surfaces.py:123-198)
 *   optk_trace_grid  optika/systems/_sequential.py:791-988 (input rays from
 *                    wavelength/field/pupil, raytrace, rayfunction)
 *   optk_eval        the per-operator protocols: sags (optika/sags/_abc.py:47-122),
 *                    apertures (optika/apertures/_apertures.py:68-102),
 *                    rulings (optika/rulings/_rulings.py:24-221),
 *                    materials (optika/materials/_materials.py:28-78,
 *                    _snells_law.py:13-366, _multilayers.py:33-532)
 *   optk_collect     optika/sensors/_sensors.py:92-171 (sensor binning)
 *   optk_batch_*     n-dimensional configurations of a system
 *                    (optika/systems/_sequential.py:2125-2132,
 *                    optika/surfaces.py:386-395), with ragged ray counts
 *   optk_stop_newton the stop solve's root finding
 *                    (optika/systems/_sequential.py:396-662)
 *   optk_pupil_reduce the pupil reductions of distortion / vignetting /
 *                    area_effective (optika/systems/_sequential.py:1208-1512)
 *   optk_denormalize the normalized grid's denormalization
 *                    (optika/systems/_sequential.py:748-789)
 *   optk_expose      AbstractImagingSensor.expose with the ideal sensor
 *                    material (optika/sensors/_sensors.py:173-252)
 */
#ifndef OPTK_H
#define OPTK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OPTK_ABI_VERSION 1

/* return codes */
#define OPTK_OK 0
#define OPTK_EINVAL -1  /* bad argument */
#define OPTK_EDEVICE -2 /* HIP runtime error */
#define OPTK_ENOMEM -3  /* device allocation failed */

/* precision of a trace */
#define OPTK_F64 0 /* float64 throughout: the drop-in form of the reference */
#define OPTK_F32 1 /* float32 compensated form (float I/O) */

#define OPTK_MAX_ORDERS 8
#define OPTK_MAX_SURFACES 64
#define OPTK_MAX_POLY 8

/* kinds -------------------------------------------------------------------*/
enum {
  OPTK_XF_IDENTITY = 0, /* no transformation */
  OPTK_XF_TRANSLATE = 1,
  OPTK_XF_AFFINE = 2
};
enum {
  OPTK_SAG_NONE = 0,
  OPTK_SAG_SPHERE = 1,
  OPTK_SAG_CONIC = 2,
  OPTK_SAG_PARABOLA = 3,
  OPTK_SAG_CYLINDER = 4,
  OPTK_SAG_TOROID = 5
};
enum {
  OPTK_MATERIAL_VACUUM = 0,
  OPTK_MATERIAL_MIRROR = 1,
  OPTK_MATERIAL_MEASURED_MIRROR = 2,
  OPTK_MATERIAL_GLASS = 3,
  OPTK_MATERIAL_MULTILAYER_MIRROR = 4,
  OPTK_MATERIAL_MULTILAYER_FILM = 5,
  /* BackIlluminatedSiliconSensorMaterial (optika/sensors/materials/
   * _materials.py:2247-2282): efficiency = the s/p average of the
   * Beer-Lambert absorbance in the silicon (:2231-2243, :200-376); stack =
   * the oxide layer over the silicon substrate layer, material_p[0] = the
   * substrate thickness (nm); index and attenuation pass through */
  OPTK_MATERIAL_SILICON_SENSOR = 6
};
enum {
  OPTK_RULINGS_NONE = 0,
  OPTK_RULINGS_IDEAL = 1,
  OPTK_RULINGS_MEASURED = 2,
  OPTK_RULINGS_SINUSOIDAL = 3,
  OPTK_RULINGS_SQUARE = 4,
  OPTK_RULINGS_SAWTOOTH = 5,
  OPTK_RULINGS_TRIANGULAR = 6,
  OPTK_RULINGS_RECTANGULAR = 7
};
enum { OPTK_SPACING_CONSTANT = 0, OPTK_SPACING_POLY1D = 1, OPTK_SPACING_HOLOGRAPHIC = 2 };
enum {
  OPTK_APERTURE_NONE = 0,
  OPTK_APERTURE_CIRCLE = 1,
  OPTK_APERTURE_SECTOR = 2,
  OPTK_APERTURE_ELLIPSE = 3,
  OPTK_APERTURE_RECTANGLE = 4,
  OPTK_APERTURE_POLYGON = 5
};
enum { OPTK_AP_ACTIVE = 1, OPTK_AP_INVERTED = 2, OPTK_AP_DIRECTION = 4 };
enum {
  OPTK_PROFILE_NONE = 0,
  OPTK_PROFILE_ERF = 1,
  OPTK_PROFILE_EXPONENTIAL = 2,
  OPTK_PROFILE_LINEAR = 3,
  OPTK_PROFILE_SINUSOIDAL = 4
};
enum { OPTK_BLOCK_SEQUENCE = 0, OPTK_BLOCK_PERIODIC = 1 };

/* records -----------------------------------------------------------------*/

/* p_parent = m * p_local + t (m row-major).  Inverse applied as m^T (p - t).
 * kind == OPTK_XF_TRANSLATE adds t only (m ignored). */
typedef struct optk_affine {
  double m[9];
  double t[3];
  int32_t kind;
  int32_t pad;
} optk_affine;

/* One optical surface: AbstractSurface (optika/surfaces.py:282-395). */
typedef struct optk_surface {
  optk_affine transform;      /* surface frame -> system frame */

  /* sag */
  int32_t sag_kind;
  int32_t pad0;
  double sag_p[4];            /* sphere: R | conic: R, k | parabola: R, -1, f |
                                 cylinder: R | toroid: R, R_rotation */
  double sag_c;               /* 1 / sag_p[0]; filled in by optk_system_create */
  optk_affine sag_transform;  /* sag frame -> surface frame */

  /* material */
  int32_t material_kind;
  int32_t is_mirror;
  double material_p[6];       /* glass: b1, b2, b3, c1, c2, c3 (c in nm^2) |
                                 silicon sensor: substrate thickness (nm) */
  int32_t material_table;     /* measured mirror: index into tables[] (wavelength
                                 row then efficiency row), -1 if none */
  int32_t material_table_len;
  int32_t stack;              /* multilayer: index into stacks[], -1 if none */
  int32_t pad1;

  /* rulings */
  int32_t rulings_kind;
  int32_t spacing_kind;
  int32_t num_orders;         /* 1, or the length of the diffraction-order axis */
  int32_t num_terms;          /* poly1d terms */
  int32_t orders[OPTK_MAX_ORDERS];
  double rulings_p[2];        /* depth (nm), duty ratio */
  double spacing_constant;    /* constant spacing */
  double wavelength_scale;    /* nm -> spacing length unit */
  double spacing_normal[3];   /* ruling-plane normal (constant, poly1d) */
  double poly_coeff[OPTK_MAX_POLY];
  int32_t poly_power[OPTK_MAX_POLY];
  double holo_x1[3];
  double holo_x2[3];
  double holo_wavelength;
  int32_t holo_diverging1;
  int32_t holo_diverging2;
  optk_affine spacing_transform; /* poly1d argument transformation */
  int32_t rulings_table;      /* measured rulings: index into tables[], -1 if none */
  int32_t rulings_table_len;

  /* aperture */
  int32_t aperture_kind;
  int32_t aperture_flags;     /* OPTK_AP_* */
  double aperture_p[4];       /* circle: r | sector: r, a0, a1 (rad) |
                                 ellipse: rx, ry | rectangle: hx, hy */
  int32_t polygon_offset;     /* polygon: index into tables[] (x row, y row) */
  int32_t polygon_count;
  optk_affine aperture_transform;
} optk_surface;

/* One homogeneous layer: Layer (optika/materials/_layers.py:134-277). */
typedef struct optk_layer {
  int32_t chemical;           /* index into chemicals[], -1 = vacuum (n = 1) */
  int32_t profile_kind;       /* interface roughness profile, OPTK_PROFILE_* */
  double thickness;           /* nm */
  double profile_width;       /* nm */
} optk_layer;

/* A run of layers: a LayerSequence, or a PeriodicLayerSequence
 * (optika/materials/_layers.py:473-499, 611-645). */
typedef struct optk_block {
  int32_t kind;               /* OPTK_BLOCK_* */
  int32_t first_layer;        /* index into layers[] */
  int32_t num_layers;
  int32_t num_periods;
} optk_block;

#define OPTK_MAX_BLOCKS 8

/* A multilayer coating: blocks applied in order, then the substrate
 * (thickness forced to zero, optika/materials/_multilayers.py:186-222). */
typedef struct optk_stack {
  int32_t num_blocks;
  int32_t substrate_layer;    /* index into layers[], -1 = vacuum */
  optk_block blocks[OPTK_MAX_BLOCKS];
} optk_stack;

/* An n/k table: rows (wavelength in Angstrom, n, k) stored as three
 * consecutive rows of `count` doubles starting at tables[offset]. */
typedef struct optk_chemical {
  int32_t offset;
  int32_t count;
} optk_chemical;

/* The whole system.  Host pointers; optk_system_create copies them. */
typedef struct optk_system_desc {
  const optk_surface *surfaces;
  int32_t num_surfaces;
  int32_t num_paths;          /* length of the diffraction-order axis (>= 1) */
  const optk_layer *layers;
  int32_t num_layers;
  int32_t num_stacks;
  const optk_stack *stacks;
  const optk_chemical *chemicals;
  int32_t num_chemicals;
  int32_t pad;
  const double *tables;
  int64_t num_tables;         /* doubles in tables[] */
  optk_affine pre[2];         /* applied to input rays in order: the object's
                                 transformation (forward) and the system's
                                 (inverse), or identity */
  int32_t pre_is_inverse[2];  /* 1: apply pre[i] as an inverse */
  optk_affine post;           /* applied (inverse) to final rays: the sensor's
                                 transformation for rayfunction, or identity */
} optk_system_desc;

typedef struct optk_system optk_system; /* opaque, device-resident */

/* Ray buffers: one device array per component.  For input, a NULL field
 * takes the matching *_default; for output, a NULL field is not written.
 * Element type is double (OPTK_F64) or float (OPTK_F32); the mask is uint8. */
typedef struct optk_rays {
  void *position[3];
  void *direction[3];
  void *wavelength;
  void *intensity;
  void *attenuation;
  void *index_refraction;
  uint8_t *unvignetted;
  double wavelength_default;
  double intensity_default;
  double attenuation_default;
  double index_default;
  int32_t unvignetted_default;
  int32_t pad;
} optk_rays;

/* Input grid of SequentialSystem (ObjectVectorArray): field and pupil are
 * angles in degrees or positions in mm depending on object_at_infinity. */
typedef struct optk_grid {
  const void *field[2];
  const void *pupil[2];
  const void *wavelength;
  const void *intensity;
  double field_default[2];
  double pupil_default[2];
  double wavelength_default;
  double intensity_default;
  int32_t object_at_infinity;
  int32_t pad;
} optk_grid;

/* flags for optk_trace / optk_trace_grid */
#define OPTK_TRACE_APPLY_POST 1 /* apply desc.post (rayfunction) */
#define OPTK_TRACE_GENERIC 4    /* use the generic kernels even where the system's
                                   surface kinds have a specialised kernel (the
                                   results are the same; for comparisons) */

int32_t optk_abi_version(void);
const char *optk_last_error(void);
/* sizeof of each record, so bindings can check their layout */
int64_t optk_sizeof(const char *record_name);

int32_t optk_system_create(const optk_system_desc *desc, optk_system **out);
int32_t optk_system_destroy(optk_system *sys);
/* The specialised kernels a system runs (no reference counterpart:
 * introspection).  The library holds trace kernels with the surface kinds
 * of some layouts fixed at compile time (the Newtonian telescope and the
 * EUV spectrograph); a system whose records carry exactly such a layout's
 * kinds runs them on the grid entry point's sensor-frame layout and the ray
 * entry point's full-output layout, with the same results as the generic
 * kernels.  Returns the layout's index, or -1 (generic kernels only). */
int32_t optk_system_specialised(const optk_system *sys);

/* Trace n input rays through every surface in one fused pass.
 * out fields have num_paths * n elements (path-major: path p of ray i at
 * p * n + i).  accum (may be NULL) receives the rays after every surface:
 * num_surfaces * num_paths * n elements, surface-major. */
int32_t optk_trace(const optk_system *sys, int32_t precision, const optk_rays *in,
                   const optk_rays *out, const optk_rays *accum, int64_t n,
                   int32_t flags, void *hip_stream);

/* As optk_trace, with input rays generated from the grid on the device
 * (optika/systems/_sequential.py:791-828). */
int32_t optk_trace_grid(const optk_system *sys, int32_t precision, const optk_grid *in,
                        const optk_rays *out, const optk_rays *accum, int64_t n,
                        int32_t flags, void *hip_stream);

/* Per-operator evaluation on surface `surface` of `sys` (float64 only).
 * in[]/out[] are device arrays of n doubles; the layout per op is: */
enum {
  OPTK_OP_SAG_VALUE = 0,     /* in: px py pz              -> out: z          */
  OPTK_OP_SAG_NORMAL = 1,    /* in: px py pz              -> out: nx ny nz   */
  OPTK_OP_SAG_INTERCEPT = 2, /* in: px py pz dx dy dz attenuation intensity
                                -> out: px py pz dx dy dz intensity (sag.propagate_rays) */
  OPTK_OP_APERTURE = 3,      /* in: px py pz              -> out_mask        */
  OPTK_OP_SNELL = 4,         /* in: dx dy dz n1 n2 nx ny nz, iparam: is_mirror -> out: bx by bz */
  OPTK_OP_MULTILAYER = 5,    /* in: wavelength cos_re cos_im n_re n_im
                                -> out: Rs Rp Ts Tp rs_re rs_im rp_re rp_im ts_re ts_im tp_re tp_im */
  OPTK_OP_SPACING = 6,       /* in: px py pz nx ny nz     -> out: kx ky kz   */
  OPTK_OP_INCIDENT = 7,      /* in: wavelength dx dy dz n nx ny nz kx ky kz (ruling vector),
                                iparam: path -> out: ex ey ez */
  OPTK_OP_RULINGS_EFF = 8,   /* in: wavelength dx dy dz px py pz nx ny nz, iparam: path -> out: eff */
  OPTK_OP_MATERIAL_EFF = 9,  /* in: wavelength dx dy dz n attenuation nx ny nz -> out: eff */
  OPTK_OP_INDEX = 10,        /* in: wavelength n                   -> out: n2 */
  OPTK_OP_SNELL_SCALAR = 11  /* in: cos_re cos_im n1_re n1_im n2_re n2_im -> out: re im */
};
typedef struct optk_eval_args {
  const double *in[12];
  double *out[12];
  uint8_t *out_mask;
  int32_t surface;
  int32_t iparam;
  int32_t stack;              /* OPTK_OP_MULTILAYER: index into stacks[] */
  int32_t pad;
} optk_eval_args;

int32_t optk_eval(const optk_system *sys, int32_t op, const optk_eval_args *args,
                  int64_t n, void *hip_stream);

/* Sensor binning (AbstractImagingSensor.collect): weighted 3-D histogram
 * over (wavelength edges, x edges, y edges) with x/y edges
 * numpy.linspace(lo, hi, num + 1), right edge of the last bin inclusive.
 * Rays with unvignetted == 0 or outside every bin are skipped.
 * image and moment are float64 device arrays of n_wavelength * nx * ny,
 * accumulated into (not cleared).  moment gets intensity * d_z.
 * Any pixel range is binned exactly by numpy's edges: where num * (|lo| +
 * |hi|) / (hi - lo) reaches 2^48 (a range narrow for its distance from the
 * origin) the bins come from a binary search over the edges, in the one-pass
 * kernel (no workspace used).
 * The workspace size depends on precision (set it before the query). */
typedef struct optk_collect_args {
  const void *position[2];
  const void *direction_z;
  const void *wavelength;
  const void *intensity;
  const uint8_t *unvignetted;
  const double *wavelength_edges; /* device, n_wavelength + 1 ascending */
  int32_t n_wavelength;
  int32_t nx;
  int32_t ny;
  int32_t precision;              /* element type of the ray arrays */
  double x_lo, x_hi, y_lo, y_hi;
  double *image;
  double *moment;
  /* optional device workspace of optk_collect_workspace() bytes: with it the
   * rays are counting-sorted into 32 x 32-pixel tiles and each tile is
   * summed in LDS (tiles holding many rays by several blocks, which add
 * with float64 atomics); without it (or when the image has
   * more than 16384 tiles, or for 2^31 rays or more) every ray adds to the
   * image with float64 global atomics */
  void *workspace;
  uint64_t workspace_bytes;
} optk_collect_args;

int32_t optk_collect(const optk_collect_args *args, int64_t n, void *hip_stream);
/* bytes of workspace the tiled binning of n rays needs (0: not applicable) */
uint64_t optk_collect_workspace(const optk_collect_args *args, int64_t n);

/* Sensor binning that allocates nothing.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as optk_collect with a workspace, except that
 * args->workspace / workspace_bytes are ignored: the tiled binning uses a
 * library-owned device array of OPTK_COLLECT_SCRATCH_BYTES, one per GPU,
 * held until the process ends.  A call whose workspace need
 * (optk_collect_workspace) exceeds it -- about 3.2M float32 or 2.9M float64
 * rays on a 2048 x 2048 image -- is rejected with OPTK_EINVAL before any
 * device work, as is one of 2^31 rays or more; images of more than 16384
 * tiles take the one-pass path, which needs no workspace.  Every call on
 * one GPU shares that GPU's array: calls that overlap on one GPU (on
 * different streams, or from host threads without ordering) race.  The
 * array keeps the last call's contents; every entry a call reads, it has
 * written first, so nothing needs clearing between calls.  No hipMalloc,
 * hipFree or allocating library routine is called. */
#define OPTK_COLLECT_SCRATCH_BYTES (64ull << 20)
int32_t optk_collect_noalloc(const optk_collect_args *args, int64_t n, void *hip_stream);

/* Exposure of a binned image by an ideal sensor material (quantum yield 1):
 * optika/sensors/_sensors.py:173-252 (expose) with
 * optika/sensors/materials/_materials.py:1576-1601 (IdealSensorMaterial.signal).
 * Per pixel: expected photons = scale * photons[i] (scale: the exposure
 * time); with noise, a Poisson draw of that mean plus zero-mean Gaussian
 * read noise of rms read_noise; without, the expectation itself.  Random
 * numbers are Philox4x32-10 (Salmon et al. 2011) keyed by seed and counted
 * by pixel, so a call is reproducible and independent of the launch shape.
 * Poisson: multiplication of uniforms below a mean of 10, Hormann's
 * transformed rejection (PTRS, 1993) above. */
typedef struct optk_expose_args {
  const double *photons;
  double *electrons;
  double scale;
  double read_noise;
  uint64_t seed;
  int32_t noise;
  int32_t pad;
} optk_expose_args;

int32_t optk_expose(const optk_expose_args *args, int64_t n, void *hip_stream);

/* Batched trace: many instances of one system layout (the same surface
 * kinds, layer stacks and tables; different parameters: tolerance
 * perturbations), each with its own number of rays, in ONE persistent
 * launch.  Replaces the reference's n-dimensional configurations, where
 * surface parameters broadcast along named axes of the ray arrays
 * (SequentialSystem.shape, optika/systems/_sequential.py:2125-2132;
 * Surface.shape, optika/surfaces.py:386-395) and every instance sees the
 * same (dense) ray grid. */
typedef struct optk_batch optk_batch;

typedef struct optk_batch_desc {
  const optk_system_desc *base;  /* layers, stacks, chemicals, tables and
                                    num_surfaces / num_paths of every instance */
  const optk_surface *surfaces;  /* num_instances * base->num_surfaces records,
                                    instance-major */
  const optk_affine *pre;        /* num_instances * 2, or NULL: base->pre */
  const optk_affine *post;       /* num_instances, or NULL: base->post */
  int32_t num_instances;
  int32_t pad;
} optk_batch_desc;

/* launch shape of the persistent kernel: threads per block, rays per thread
 * per work tile (a tile is block * rays_per_thread rays of one instance),
 * blocks resident per CU (0: as many as fit) */
typedef struct optk_tune {
  int32_t block;
  int32_t rays_per_thread;
  int32_t blocks_per_cu;
  int32_t pad;
  double ms; /* measured time of one batched trace (optk_batch_autotune) */
} optk_tune;

/* per (instance, path): unvignetted count, sum of intensity, sums of x, y,
 * x^2, y^2 of the final positions (unvignetted rays), RMS spot radius about
 * the centroid (mm), throughput = sum of intensity / rays traced */
#define OPTK_BATCH_STATS 8
#define OPTK_BATCH_GRAPH 2 /* flag: capture the launch in a hipGraph once,
                              replay it while the arguments are unchanged */

int32_t optk_batch_create(const optk_batch_desc *desc, optk_batch **out);
int32_t optk_batch_destroy(optk_batch *batch);

/* Trace the rays of every instance: instance i owns grid rays
 * offsets[i] .. offsets[i+1]-1 (offsets: host array of num_instances + 1,
 * offsets[0] = 0).  out (device rays, num_paths * n, path-major) may be
 * NULL; stats (device, num_instances * num_paths * OPTK_BATCH_STATS
 * doubles) is overwritten.  tune NULL: the last autotuned shape, else a
 * default.  flags: OPTK_TRACE_APPLY_POST, OPTK_BATCH_GRAPH,
 * OPTK_TRACE_GENERIC (the generic kernels even where every instance matches
 * a specialised layout).  One call at a time per batch (the work-queue
 * counter is the batch's). */
int32_t optk_batch_trace(optk_batch *batch, int32_t precision, const optk_grid *in,
                         const int64_t *offsets, const optk_rays *out, double *stats,
                         const optk_tune *tune, int32_t flags, void *hip_stream);

/* Time every candidate launch shape on this batch and these inputs (each
 * run is a full optk_batch_trace without graph capture); return the
 * fastest, which later calls with tune == NULL use. */
int32_t optk_batch_autotune(optk_batch *batch, int32_t precision, const optk_grid *in,
                            const int64_t *offsets, double *stats, int32_t flags,
                            void *hip_stream, optk_tune *best);

/* Learned surrogate (no counterpart in the reference: it stands in for the
 * trace of one system, which produces its training data and judges it).
 * Kernels for a 5 -> 256 -> 256 -> 256 -> 256 -> 3 SiLU MLP with bf16
 * weights/activations and float32 accumulation on MFMA; the host side
 * (optika_amd/surrogate.py) sequences them. */

/* C[m][n] = sum_k A(m, k) B(k, n); A, B bf16 (uint16 bits), C float32.
 * A(m, k) = a[m * lda + k] (trans_a = 0) or a[k * lda + m] (trans_a = 1);
 * B(k, n) = b[k * ldb + n] (trans_b = 0) or b[n * ldb + k] (trans_b = 1).
 * m, n: multiples of 64; k: a multiple of 32 * split_k.  split_k > 1 sums
 * partial products over k in `workspace` (split_k * m * n floats, ldc == n). */
int32_t optk_gemm_bf16(int32_t trans_a, int32_t trans_b, int64_t m, int64_t n, int64_t k,
                       const void *a, int64_t lda, const void *b, int64_t ldb, float *c,
                       int64_t ldc, int32_t split_k, float *workspace, void *hip_stream);

/* optk_gemm_bf16 (no split) with the layer's elementwise step fused into
 * its epilogue, so the result is not re-read by a separate kernel:
 *   OPTK_GEMM_BIAS_SILU  c[i][j] = (A B)[i][j] + bias[j] (the pre-activation,
 *                        float32) and out16[i][j] = bf16(silu(c[i][j]));
 *   OPTK_GEMM_SILU_BACK  c[i][j] = (A B)[i][j] * silu'(pre[i][j]) (float32;
 *                        c may be NULL: not stored) and
 *                        out16[i][j] = bf16((A B)[i][j] * silu'(pre[i][j]));
 * silu(x) = x / (1 + e^-x).  c, pre and out16 are [m][ldc].  m and n
 * multiples of 128, k a multiple of 32, lda and ldb multiples of 8, ldc a
 * multiple of 4; a, b, c, pre and bias 16-byte aligned, out16 8-byte.
 * With trans_a = 0, trans_b = 1 and n = k = 256 (the surrogate's hidden
 * layers) the product runs as row panels that read each row of A once;
 * the results are the same bits. */
#define OPTK_GEMM_BIAS_SILU 1
#define OPTK_GEMM_SILU_BACK 2
int32_t optk_gemm_bf16_fused(int32_t trans_a, int32_t trans_b, int64_t m, int64_t n, int64_t k,
                             const void *a, int64_t lda, const void *b, int64_t ldb, float *c,
                             int64_t ldc, int32_t epilogue, const float *bias, const float *pre,
                             void *out16, void *hip_stream);

/* optk_gemm_bf16_fused's OPTK_GEMM_SILU_BACK for the surrogate's 256-wide
 * layers (trans_a = 0, trans_b = 1, n = k = 256) that also sums the float32
 * result over the rows -- the previous layer's bias gradient:
 *   d[i][j] = (A B)[i][j] * silu'(pre[i][j]), A(i, k) = a[i * lda + k],
 *   B(k, j) = b[j * ldb + k];
 *   out16[i][j] = bf16(d[i][j]); c[i][j] = d[i][j] (c may be NULL: not
 *   stored); sums[j] = sum_i d[i][j] (float32, added in a fixed order).
 * m a positive multiple of 64; pre, c and out16 are [m][ldc]; workspace
 * holds (m / 64) * 256 floats of per-tile partial sums.  lda, ldb >= 256
 * and multiples of 8, ldc >= 256 and a multiple of 4; a, b, pre, c and
 * workspace 16-byte aligned, out16 8-byte. */
int32_t optk_gemm_bf16_silu_back_sums(int64_t m, const void *a, int64_t lda, const void *b, int64_t ldb,
                                      const float *pre, int64_t ldc, float *c, void *out16, float *sums,
                                      float *workspace, void *hip_stream);

/* elementwise and small-layer pieces of the forward and backward passes
 * (argument use per op below; rows x cols row-major arrays) */
enum {
  OPTK_MLP_BIAS_SILU = 0,     /* f32[0] pre (+= bias, in place), f32[1] bias; u16[0] act = silu */
  OPTK_MLP_DENSE_IN = 1,      /* f32[0] x [rows][inner], f32[1] w [cols][inner], f32[2] b,
                                 f32[3] pre (or NULL); u16[0] act */
  OPTK_MLP_DENSE_OUT = 2,     /* u16[0] act [rows][cols], f32[0] w [inner][cols], f32[1] b,
                                 f32[2] y [rows][inner] */
  OPTK_MLP_MSE_GRAD = 3,      /* f32[0] y, f32[1] target, f32[2] weight per column, f32[3] dy,
                                 f32[4] loss partial per 256 elements */
  OPTK_MLP_DENSE_OUT_BACK = 4,/* f32[0] dy [rows][inner], f32[1] w [inner][cols], f32[2] pre;
                                 u16[0] dpre (bf16), f32[3] dpre (float32, or NULL), f32[4]
                                 column sums of dpre (the bias gradient, or NULL; cols = 256,
                                 inner = 3), f32[5] their workspace of count floats (at least
                                 min(1024, ceil(rows / 4)) * 256) */
  OPTK_MLP_SILU_BACK = 5,     /* f32[0] dact, f32[1] pre; u16[0] dpre, f32[2] dpre (float32) */
  OPTK_MLP_OUTER_GRAD = 6,    /* out = sum_r x[r][c] g[r][j]: f32[0] x (or u16[0] bf16 x),
                                 f32[1] g [rows][inner], f32[2] out ([inner][cols], or
                                 [cols][inner] with flag = 1), f32[3] workspace of count floats */
  OPTK_MLP_COLUMN_SUM = 7,    /* f32[0] g [rows][cols]; f32[1] sums [cols] */
  OPTK_MLP_ADAM = 8,          /* f32[0] params, f32[1] grads, f32[2] m, f32[3] v (count each);
                                 u16[0] bf16 copy of params (or NULL);
                                 scalars: lr, beta1, beta2, eps, 1 - beta1^t, 1 - beta2^t */
  OPTK_MLP_TO_BF16 = 9,       /* f32[0] x -> u16[0] (rows * cols) */
  OPTK_MLP_TRANSPOSE16 = 10   /* u16[0] x [rows][cols] -> u16[1] y [cols][rows] (distinct) */
};
typedef struct optk_mlp_args {
  float *f32[8];
  uint16_t *u16[4];
  int64_t rows;
  int64_t count;
  int32_t cols;
  int32_t inner;
  int32_t flag;
  int32_t pad;
  float scalars[8];
} optk_mlp_args;

int32_t optk_mlp_op(int32_t op, const optk_mlp_args *args, void *hip_stream);

/* Fused inference of a trained surrogate.  Per ray: inputs normalised as
 * (v - in_mid) * in_inv_half; the network's three outputs scaled by
 * `scale` and added to a cubic polynomial of the normalised inputs
 * (OPTK_SURROGATE_POLY monomials; poly_exp packs each monomial's exponents,
 * 2 bits per input; poly[t][j] its coefficient for target j).  Targets:
 * sensor x, y, log intensity, then (x, y) in the frame of every listed
 * aperture.  A ray falls back to the trace when it is predicted within
 * edge_margin of an aperture edge, its input cell (16 wavelength x 8 field
 * x x 8 field y bins) is flagged in error_map, or a prediction is not
 * finite; fallback rays are appended to fallback_list (one atomic per
 * wavefront; order unspecified; entries past fallback_capacity are
 * counted but not written). */
#define OPTK_SURROGATE_POLY 56
#define OPTK_SURROGATE_MAX_APERTURES 4
#define OPTK_SURROGATE_MAX_TARGETS (3 + 2 * OPTK_SURROGATE_MAX_APERTURES)
#define OPTK_SURROGATE_MAX_VERTICES 16
typedef struct optk_surrogate_aperture {
  int32_t kind;                 /* OPTK_APERTURE_CIRCLE, _RECTANGLE or _POLYGON */
  int32_t inverted;
  int32_t num_vertices;
  int32_t pad;
  float p[4];                   /* circle: r; rectangle: half widths */
  float vx[OPTK_SURROGATE_MAX_VERTICES];
  float vy[OPTK_SURROGATE_MAX_VERTICES];
} optk_surrogate_aperture;
typedef struct optk_surrogate_desc {
  const float *w1, *b1;         /* [256][5], [256] float32 */
  const uint16_t *w1_16;        /* bf16 [256][32]: row o = [hi(w1[o]) | hi(w1[o]) |
                                   lo(w1[o]) | 0], lo = bf16(w - hi) */
  const uint16_t *w16[3];       /* hidden layers, bf16 [256 out][256 in] */
  const float *b[3];            /* hidden biases [256] */
  const float *w5, *b5;         /* [3][256], [3] float32 */
  const uint16_t *w5_16;        /* bf16 [16][256]: w5 = rows 0..2 + rows 3..5
                                   (high and low parts), zeros below */
  const float *poly;            /* [OPTK_SURROGATE_POLY][OPTK_SURROGATE_MAX_TARGETS] */
  const uint8_t *error_map;     /* 16 * 8 * 8 flags, or NULL */
  float in_mid[5], in_inv_half[5];
  float scale[3];
  float edge_margin;
  uint32_t poly_exp[OPTK_SURROGATE_POLY];
  int32_t num_targets;
  int32_t num_apertures;
  optk_surrogate_aperture apertures[OPTK_SURROGATE_MAX_APERTURES];
} optk_surrogate_desc;
typedef struct optk_surrogate_io {
  const float *field_x, *field_y, *pupil_x, *pupil_y, *wavelength;
  float *x, *y, *intensity;
  uint8_t *unvignetted, *fallback;
  int32_t *fallback_list;       /* or NULL */
  int32_t *fallback_count;      /* device counter, incremented (not cleared) */
  int64_t fallback_capacity;
  int64_t n;
} optk_surrogate_io;

int32_t optk_surrogate_infer(const optk_surrogate_desc *desc, const optk_surrogate_io *io,
                             void *hip_stream);

/* The stop solve's root finding (optika/systems/_sequential.py:396-662): for
 * each of n rays, the 2-vector x -- the free direction (x, y, with
 * z = sqrt(1 - x^2 - y^2)) or the free position (x, y, with z the sag of
 * surface 0 of `first` at (x, y, 0)) -- that carries the ray through every
 * surface of `sys` onto target[i]: the position (x, y) or direction (x, y)
 * in the frame of sys's post transformation (applied inversely, as
 * OPTK_TRACE_APPLY_POST does).  Newton steps x -= J^-1 f with a forward
 * difference Jacobian of step dx, per ray, until both residual components
 * are below tol, at most max_iter steps; the reference iterates
 * named_arrays.optimize.root_newton with named_arrays.jacobian the same way.
 * One thread per ray, the residual and both Jacobian columns traced in
 * registers.  x: initial guess in, root out; iterations[i]: steps taken, or
 * max_iter + 1 when the ray did not converge.  Only the first diffraction
 * order path is followed. */
#define OPTK_STOP_FREE_DIRECTION 0
#define OPTK_STOP_FREE_POSITION 1
#define OPTK_STOP_TARGET_POSITION 0
#define OPTK_STOP_TARGET_DIRECTION 1
typedef struct optk_stop_args {
  const double *position[3];    /* fixed start position (system frame) */
  const double *direction[3];   /* fixed start direction */
  const double *wavelength;
  const double *target[2];
  double *x[2];
  int32_t *iterations;
  double *z;                    /* optional (NULL: not written): the free
                                   component's third coordinate at the root,
                                   sqrt(1 - x0^2 - x1^2) for a direction, the
                                   first surface's sag at (x0, x1) for a
                                   position */
  double tol;
  double dx;
  int32_t free_component;       /* OPTK_STOP_FREE_* */
  int32_t target_component;     /* OPTK_STOP_TARGET_* */
  int32_t max_iter;
  int32_t pad;
} optk_stop_args;

int32_t optk_stop_newton(const optk_system *sys, const optk_system *first,
                         const optk_stop_args *args, int64_t n, void *hip_stream);

/* Reductions of traced rays over the pupil (optika/systems/_sequential.py:
 * 1208-1512: distortion, vignetting, area_effective): the rays of segment s
 * are rays [s * per_segment, (s + 1) * per_segment).  Per segment:
 *   count       number of unvignetted rays
 *   sum_x/_y    position sums over the unvignetted rays
 *   all_x/_y    position sums over every ray
 *   throughput  sum of weight[p] * intensity over the unvignetted rays
 *               (weight: one per pupil sample p, NULL = 1; intensity NULL = 1)
 * Sums run in a fixed order (deterministic).  Any output may be NULL. */
typedef struct optk_pupil_args {
  const double *x, *y;
  const double *intensity;
  const uint8_t *unvignetted;
  const double *weight;
  int64_t per_segment;
  double *count;
  double *sum_x, *sum_y;
  double *all_x, *all_y;
  double *throughput;
} optk_pupil_args;

int32_t optk_pupil_reduce(const optk_pupil_args *args, int64_t segments, void *hip_stream);

/* Normalized field and pupil coordinates onto the extent the stop rays span
 * (optika/systems/_sequential.py:748-789, _denormalize_grid): for each of
 * up to 4 components c (the first NULL in[c] ends the list) and each of the
 * n values,
 *   out[c][i] = ptp[c][w] * (in[c][i] + 1) / 2 + lo[c][w],
 * in float64 in the reference's operation order, w = table_index[i]
 * (table_index NULL: w = 0; entries outside [0, n_table) are clamped into
 * it).  lo and ptp are device arrays of n_table entries (one per distinct
 * wavelength).  out[c] may equal in[c].  One launch for all components. */
typedef struct optk_denormalize_args {
  const double *in[4];
  double *out[4];
  const double *lo[4];
  const double *ptp[4];
  const int64_t *table_index;
  int64_t n_table;
} optk_denormalize_args;

int32_t optk_denormalize(const optk_denormalize_args *args, int64_t n, void *hip_stream);

/* The surrogate's polynomial baseline in float64 (the fit of the targets on
 * the OPTK_SURROGATE_POLY cubic monomials F of 5 normalised inputs).
 * u: [n][5] row-major; targets: [n][num_targets] (num_targets <=
 * OPTK_SURROGATE_MAX_TARGETS); valid: n flags, rays flagged 0 add nothing;
 * poly_exp: host array of OPTK_SURROGATE_POLY packed exponents (2 bits per
 * input).  optk_poly_normal writes the normal equations
 *   out[i][j] = sum_r F[r][i] G[r][j],  G = [F | targets],
 * an [OPTK_SURROGATE_POLY][OPTK_SURROGATE_POLY + num_targets] matrix summed
 * in a fixed order (deterministic), using optk_poly_workspace(n) bytes of
 * caller-owned device workspace.  optk_poly_eval writes
 * out[n][num_targets] = F coef, coef [OPTK_SURROGATE_POLY][num_targets]. */
size_t optk_poly_workspace(int64_t n);
int32_t optk_poly_normal(const double *u, const double *targets, const uint8_t *valid, int64_t n,
                         int32_t num_targets, const uint32_t *poly_exp, double *out,
                         void *workspace, size_t workspace_bytes, void *hip_stream);
int32_t optk_poly_eval(const double *u, const double *coef, int64_t n, int32_t num_targets,
                       const uint32_t *poly_exp, double *out, void *hip_stream);

#ifdef __cplusplus
}
#endif

#endif /* OPTK_H */
