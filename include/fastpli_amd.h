/*
 * fastpli_amd -- C ABI of the MI355X (gfx950) simulation hot path.
 *
 * This library replaces the two pybind11 modules of fastpli that sit between
 * the Python `Simpli` class and its OpenMP C++ core:
 *
 *   __generation._Generator   src/simulation/bindings/generator_module.cpp:26-139
 *   __simulation._Simulator   src/simulation/bindings/simulator_module.cpp:21-94
 *
 * Conventions
 *   - Bulk arrays (fiber points, label field, vector field, images) are DEVICE
 *     pointers in HBM; small tables (layers, properties, tilts, rotations) are
 *     host pointers and are copied by the call.
 *   - Arrays are C-contiguous in the reference's layout: label[x][y][z]
 *     (z fastest), vector[x][y][z][3], images[tilt][x][y][rho].
 *   - `stream` is a hipStream_t (NULL = the null stream).  Every call is
 *     stream-ordered and returns after the result is complete on that stream.
 *   - Return value: FPL_OK (0) or a negative FPL_E* code; `msg` (may be NULL)
 *     receives the reference's error text.  The process is never aborted
 *     (the reference calls exit(), simulator.cpp:62-69).
 *   - `warn` (may be NULL) receives a bit mask of FPL_WARN_* that the caller
 *     raises as Python UserWarnings, once per kind and call (and, from the
 *     march, FPL_FLAG_NEGATIVE).
 */
#ifndef FASTPLI_AMD_H_
#define FASTPLI_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status ---- */
#define FPL_OK 0
#define FPL_E_INVALID (-1)   /* std::invalid_argument in the reference        */
#define FPL_E_LABEL_NEG (-2) /* "label < 0 detected" (simulator.cpp:159-162)   */
#define FPL_E_LABEL_MAX (-3) /* "label exceed properties.size()" (:164-167)   */
#define FPL_E_BG_DN (-4)     /* "background birefringence has to be 0" (:169) */
#define FPL_E_DEVICE (-5)    /* HIP runtime failure or no device              */
#define FPL_E_NOMEM (-6)     /* device allocation failed                      */
#define FPL_E_PENDING (-7)   /* fpl_server_wait: not finished within timeout */

/* ----------------------------------------------------------- warnings --- */
#define FPL_WARN_OVERLAP (1u << 0)      /* "objects overlap" generator.cpp:357      */
#define FPL_WARN_SINGLE_POINT (1u << 1) /* "fiber object with only single point" :215 */
#define FPL_WARN_ZERO_VEC (1u << 2)     /* "optical_axis is 0 for tissue > 0" simulator.cpp:294 */
#define FPL_WARN_NOT_NORM (1u << 3)     /* "optical_axis not normalized" simulator.cpp:300 */
#define FPL_WARN_NO_LIGHT (1u << 4)     /* "light_positions is empty" simulator.cpp:715 */
/* Not a warning: the march wrote an image value < 0.  The reference checks the
 * images after the native call and raises ValueError("intensity < 0 detected")
 * (simpli.py:184-185); the Python layer raises it for this bit. */
#define FPL_FLAG_NEGATIVE (1u << 8)

/* -------------------------------------------------------------- types --- */

/* Volume: PliGenerator::SetVolume(global_dim, origin, voxel_size),
 * generator.cpp:45-124 (bound at generator_module.cpp:28-35). */
typedef struct fpl_volume {
   int64_t dim[3];    /* voxels, each > 0                                   */
   double origin[3];  /* micro meter                                        */
   double voxel_size; /* micro meter, > 0                                   */
} fpl_volume;

/* Fiber bundles: PliGenerator::SetFiberBundles + fiber::Bundle(...),
 * generator_module.cpp:36-66, fiber_bundle.cpp:76-92.  All fibers of all
 * bundles are packed into one (n_points, 4) array (x, y, z, r) in
 * bundle -> fiber -> point order. */
typedef struct fpl_fiber_bundles {
   int64_t n_bundles;
   int64_t n_fibers;
   int64_t n_points;
   const double *points;        /* DEVICE (n_points, 4) float64             */
   const int64_t *fiber_ptr;    /* DEVICE (n_fibers + 1) point offsets      */
   const int64_t *bundle_ptr;   /* HOST (n_bundles + 1) fiber offsets       */
   const int64_t *layer_ptr;    /* HOST (n_bundles + 1) offsets into layers */
   const double *layer_scale;   /* HOST layer scale (any order per bundle)  */
   const double *layer_dn;      /* HOST birefringence                       */
   const double *layer_mu;      /* HOST absorption [1/mm]                   */
   const uint8_t *layer_model;  /* HOST 'b', 'p' or 'r'                     */
} fpl_fiber_bundles;

/* Optical setup: setup::Setup, setup.hpp:47-82 (bound at
 * simulator_module.cpp:23-54). */
typedef struct fpl_setup {
   double step_size;
   double light_intensity;
   double voxel_size;
   double wavelength;          /* nano meter */
   double tissue_refrection;
   int32_t interpolate;        /* 0 NN, 1 Lerp, 2 Slerp */
   int32_t untilt_sensor_view; /* bool */
   int32_t flip_z;             /* bool */
   int32_t n_rho;
   const double *filter_rotations; /* HOST (n_rho) radian */
   /* Storage of the fields inside the march (no counterpart in the
    * reference, which reads int32 / float32):
    *   0 f32      int32 label + float32 axis, 16 B/voxel (exact input)
    *   1 f16      uint16 label + fp16 axis, 8 B/voxel
    *   2 angle16  uint16 label + 16-bit polar/azimuth codes, 6 B/voxel
    * Modes 1 and 2 need labels < 65536 / 32768 and treat axes that pass the
    * reference's normalisation test as unit vectors. */
   int32_t field_storage;
} fpl_setup;

/* ---------------------------------------------------------- entry points */

/* Number of rows of the property table fpl_generate writes:
 * n_bundles * max_layer + 1 (generator.cpp:496-500). */
int64_t fpl_num_properties(const fpl_fiber_bundles *fbs);

/* Tissue generation.  Replaces _Generator.set_volume + set_fiber_bundles +
 * run_generation (generator_module.cpp:28-66,98-134) ->
 * PliGenerator::RunTissueGeneration (generator.cpp:165-261).
 *   label  DEVICE int32 (dim[0], dim[1], dim[2]), fully written
 *   vector DEVICE float (dim..., 3), fully written; ignored if only_label
 *   props  HOST float64 (fpl_num_properties(fbs), 2) = (dn, mu) per label
 * Labels are bit-identical to the reference's FP64 voxelizer. */
int fpl_generate(const fpl_volume *vol, const fpl_fiber_bundles *fbs,
                 int only_label, int32_t *label, float *vector, double *props,
                 uint32_t *warn, void *stream, char *msg, size_t msg_len);

/* Ray march of one or more tilts in one call.  Replaces
 * _Simulator.set_pli_setup + run_simulation (simulator_module.cpp:23-93) ->
 * PliSimulator::RunSimulation (simulator.cpp:182-411), called once per tilt
 * by the reference.
 *   label  DEVICE int32 (dim...), vector DEVICE float (dim..., 3)
 *   props  HOST float64 (n_props, 2)
 *   tilts  HOST float64 (n_tilts, 2) = (theta, phi) radian
 *   images DEVICE float64 (n_tilts, dim[0], dim[1], n_rho)
 * Intensities are computed in FP32 with compensated accumulation. */
int fpl_simulate(const fpl_setup *setup, const int64_t dim[3],
                 const int32_t *label, const float *vector,
                 const double *props, int64_t n_props, const double *tilts,
                 int64_t n_tilts, double *images, uint32_t *warn, void *stream,
                 char *msg, size_t msg_len);

/* ------------------------------------------------------- march fields ---
 * The label and optical-axis fields in the march's own layout: [x][z][y]
 * records, so that the rays of a wave (adjacent y) read one contiguous run per
 * z-step.  No counterpart in the reference, whose simulator reads the C-order
 * arrays of run_simulation (simulator_module.cpp:56-93); this handle lets a
 * caller that generates a tissue and then simulates it skip the repacking pass
 * (fpl_generate_fields writes the records from the voxelizer), and lets a
 * caller that simulates one tissue many times pack it once (fpl_pack_fields).
 *
 * Storage (field_storage):
 *   0 f32      int32 label + float32 axis, 16 B/voxel (exact input)
 *   1 f16      uint16 label + fp16 axis, 8 B/voxel
 *   2 angle16  uint16 label + 16-bit polar/azimuth codes, 2 + 4 B/voxel
 * The caller allocates the arrays: `records` takes fpl_record_bytes(storage)
 * bytes per voxel, `codes` (angle16 only) 4 bytes per voxel.
 * fpl_fields_bytes(dim, storage) is the size of one allocation holding both,
 * `codes` at byte offset fpl_fields_codes_offset(dim, storage). */
typedef struct fpl_fields {
   int64_t dim[3];
   int32_t storage;
   int32_t label_min; /* label range of the fields: set by the call that fills */
   int32_t label_max; /* them, checked against the property table by          */
   int32_t pad_;      /* fpl_simulate_fields (CheckInput, simulator.cpp:149-167) */
   void *records;     /* DEVICE */
   uint32_t *codes;   /* DEVICE, angle16 only, else NULL */
} fpl_fields;

size_t fpl_record_bytes(int32_t field_storage);
size_t fpl_fields_bytes(const int64_t dim[3], int32_t field_storage);
size_t fpl_fields_codes_offset(const int64_t dim[3], int32_t field_storage);

/* Tissue generation straight into march fields: fpl_generate whose voxelizer
 * also writes every voxel's record (same labels and axes, bit for bit).
 * label / vector (DEVICE, C order) may be NULL: then only the records are
 * written.  fields->records / codes / storage are inputs; dim and the label
 * range are set.  Labels must fit the storage (n_props <= 65535 for f16,
 * <= 32767 for angle16). */
int fpl_generate_fields(const fpl_volume *vol, const fpl_fiber_bundles *fbs,
                        int32_t *label, float *vector, double *props, fpl_fields *fields,
                        uint32_t *warn, void *stream, char *msg, size_t msg_len);

/* Packs caller-owned C-order fields (label DEVICE int32 (dim...), vector
 * DEVICE float (dim..., 3)) into march fields and records their label range. */
int fpl_pack_fields(const int64_t dim[3], const int32_t *label, const float *vector,
                    fpl_fields *fields, void *stream, char *msg, size_t msg_len);

/* fpl_simulate on march fields: the same checks (the label range against the
 * property table before the march), the same kernels and results; the
 * setup's field_storage is ignored (the fields' storage applies). */
int fpl_simulate_fields(const fpl_setup *setup, const fpl_fields *fields,
                        const double *props, int64_t n_props, const double *tilts,
                        int64_t n_tilts, double *images, uint32_t *warn, void *stream,
                        char *msg, size_t msg_len);

/* fpl_simulate that allocates nothing: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * checks, kernels and results as fpl_simulate; its scratch (the packed fields
 * and the parameter block) is a static device array of 64 MiB + 64 KiB that
 * every device the library runs on holds until the process ends.  Volumes
 * whose packed fields exceed 64 MiB (4 Mi voxels in f32, 8 Mi in f16, 11 Mi in
 * angle16) or parameter blocks above 64 KiB are rejected with FPL_E_INVALID.
 * All calls on one device share the array: calls that overlap on a device
 * race and must be serialised by the caller. */
int fpl_simulate_noalloc(const fpl_setup *setup, const int64_t dim[3],
                         const int32_t *label, const float *vector,
                         const double *props, int64_t n_props, const double *tilts,
                         int64_t n_tilts, double *images, uint32_t *warn, void *stream,
                         char *msg, size_t msg_len);

/* Reverse mode of fpl_simulate (no counterpart in the reference, which has no
 * derivatives): for L = sum(grad_images * images), with images what
 * fpl_simulate returns for the same arguments,
 *   grad_images  DEVICE float64 (n_tilts, dim[0], dim[1], n_rho) = dL/dimages
 *   grad_props   HOST float64 (n_props, 2) receives (dL/d dn, dL/d mu) per label;
 *                labels on the reference's attenuation-only branch (dn == 0 or
 *                label 0, simulator.cpp:280-288) do not depend on dn: 0 there
 *   grad_vector  DEVICE float (dim..., 3) receives dL/d optical_axis, or NULL;
 *                needs setup->interpolate == 0 (NN)
 * The setup's field_storage must be 0 (f32).  The forward march is recomputed
 * with a checkpoint every 8 steps of each ray; FP32 with FP64 sums per label. */
int fpl_simulate_adjoint(const fpl_setup *setup, const int64_t dim[3],
                         const int32_t *label, const float *vector,
                         const double *props, int64_t n_props, const double *tilts,
                         int64_t n_tilts, const double *grad_images, double *grad_props,
                         float *grad_vector, void *stream, char *msg, size_t msg_len);

/* Optical resampling of an image stack, the step after run_simulation in the
 * reference's pipeline: optic.filter_resample applied per rotation image
 * (src/fastpli/simulation/optic.py:169-187, called from
 * Simpli.apply_optic_resample, src/fastpli/simulation/__simpli.py:491-549).
 * Gaussian blur with standard deviation `sigma` in source pixels (scipy
 * gaussian_filter: radius int(4 sigma + 1/2), reflect boundary; sigma = 0 skips
 * it), then box resampling by `scale` (optic.py:85-120).
 *   in  DEVICE float64 (nx, ny, n_rho)
 *   out DEVICE float64 (round(nx * scale), round(ny * scale), n_rho) */
int fpl_optic_resample(const double *in, int64_t nx, int64_t ny, int64_t n_rho, double sigma,
                       double scale, double *out, void *stream, char *msg, size_t msg_len);

/* EPA: transmittance, direction and retardation of n_pix pixels with n_rho
 * equidistant rotations in [0, pi) (src/fastpli/analysis/epa.py:9-45, float64
 * branch).  in DEVICE float64 (n_pix, n_rho); outputs DEVICE float64 (n_pix). */
int fpl_epa(const double *in, int64_t n_pix, int64_t n_rho, double *transmittance,
            double *direction, double *retardation, void *stream, char *msg, size_t msg_len);

/* ROFL tilt analysis per pixel (src/fastpli/analysis/rofl.py:17-60 with the
 * local fit of _ROFL_with_jacobi.py:403-522), as Simpli.apply_rofl calls it
 * (src/fastpli/simulation/__simpli.py:623-706) for every pixel.
 *   in   DEVICE float64 (5, n_pix, n_rho): the tilts (0, 0), (tau, 0),
 *        (tau, pi/2), (tau, pi), (tau, 3 pi/2), n_rho equidistant in [0, pi)
 *   grad_mode non-zero: the fit uses the reference's analytic Jacobian
 *        (rofl's grad_mode, MINPACK lmder); zero: finite differences (lmdif)
 *   mask DEVICE uint8 (n_pix) or NULL; masked-out pixels get NaN
 *   out  DEVICE float64 (8, n_pix): direction, inclination, t_rel, their
 *        deviations, chi^2, model evaluations */
int fpl_rofl(const double *in, int64_t n_pix, int64_t n_rho, double tilt_angle, double gain,
             double dir_offset, int grad_mode, const uint8_t *mask, double *out, void *stream,
             char *msg, size_t msg_len);

/* Exact count of Mueller voxel-steps of one tilt: n_rho x rays x z-march
 * iterations (simulator.cpp:259-261).  Host only.  -1 for a tilt whose rays
 * need more than 2^24 z-steps, which fpl_simulate rejects. */
int64_t fpl_count_voxel_steps(const fpl_setup *setup, const int64_t dim[3],
                              double theta, double phi);

/* Kernel timing with HIP events recorded on each kernel's launch stream.
 * fpl_profile_enable(1) clears and starts collection, (0) stops it.
 * fpl_profile_read(name) returns the summed milliseconds and launch count of
 * the kernel `name` ("voxelize", "march", ...) since the last enable; returns
 * 0 when the kernel was seen, -1 otherwise.  No counterpart in the reference
 * (it has no profiling); used by bench.py for the roofline figure. */
void fpl_profile_enable(int enable);
int fpl_profile_read(const char *name, double *total_ms, int64_t *launches);

/* ------------------------------------------------------------ serving
 * A device-resident request loop for many simulations at once.  No
 * counterpart in the reference (it runs one simulation per process call, MPI
 * ranks sharing one volume); it serves callers that stream simulation
 * requests of different volume sizes, fiber models, tilt sets and layer
 * parameters.
 *
 * Two kinds of request:
 *   fpl_request        the march of a registered tissue (its fields packed
 *                      once by fpl_server_add_tissue and resident) under the
 *                      request's setup, property table and tilt list;
 *   fpl_model_request  a fiber model: its tissue is generated in the request's
 *                      slot straight into march fields (fpl_generate_fields),
 *                      then marched: Simpli.generate_tissue + run_simulation of
 *                      every tilt (generator.cpp:165-261, simulator.cpp:182-411).
 * The server runs max_inflight slots, each with its own stream, device buffers
 * (grown to the largest request the slot ran) and host thread.  Pending
 * requests are admitted in arrival order: a free slot takes the oldest one when
 * the registered tissues, the memory all slots hold and what the slot must grow
 * by fit the HBM budget; otherwise the request waits until a slot finishes.
 * When every slot is idle and the request still does not fit, the slots
 * release their buffers; a request that exceeds the budget on its own fails
 * with FPL_E_NOMEM.  The march of a request is the replay of a hipGraph
 * (parameter upload, flag reset, march, warning readback) captured once per
 * slot and request shape, followed by a copy of the slot's images into the
 * request's buffer.  Requests in different slots overlap on the device; results
 * are identical to fpl_simulate's / fpl_generate + fpl_simulate's for the same
 * inputs.
 *
 * A request's work starts after the work queued on `stream` (the caller's
 * stream, may be NULL) at submit, so its output buffer may still be in use
 * there.  The caller owns the output buffer, which must stay valid until
 * fpl_server_wait has returned for the ticket.  All functions are thread safe. */
typedef struct fpl_server fpl_server;

typedef struct fpl_request {
   int64_t tissue;          /* id returned by fpl_server_add_tissue */
   const fpl_setup *setup;  /* HOST; copied at submit */
   const double *props;     /* HOST float64 (n_props, 2); copied at submit */
   int64_t n_props;
   const double *tilts;     /* HOST float64 (n_tilts, 2) = (theta, phi); copied */
   int64_t n_tilts;
   double *images;          /* DEVICE float64 (n_tilts, nx, ny, n_rho) */
   void *stream;            /* the caller's hipStream_t (see above) */
} fpl_request;

typedef struct fpl_model_request {
   fpl_volume volume;
   /* the fiber model as fpl_generate takes it, except that points and
    * fiber_ptr are HOST arrays here, which the server reads until
    * fpl_server_wait has returned for the ticket (pinned memory copies
    * fastest); the layer tables are copied at submit */
   fpl_fiber_bundles bundles;
   double background_absorption; /* row 0 of the property table (Simpli sets it) */
   const fpl_setup *setup;  /* HOST; copied; its field_storage is the fields' */
   const double *tilts;     /* HOST float64 (n_tilts, 2); copied */
   int64_t n_tilts;
   double *images;          /* DEVICE float64 (n_tilts, dim[0], dim[1], n_rho) */
   void *stream;            /* the caller's hipStream_t (see above) */
} fpl_model_request;

/* hbm_budget in bytes (<= 0: 90% of the device's free memory at creation);
 * max_inflight concurrent requests (stream slots, 1..16). */
int fpl_server_create(int device, int64_t hbm_budget, int32_t max_inflight, fpl_server **server,
                      char *msg, size_t msg_len);
/* label DEVICE int32 (dim...), vector DEVICE float (dim..., 3); packed and
 * range-checked before returning, after which the caller may free them. */
int fpl_server_add_tissue(fpl_server *server, const int64_t dim[3], const int32_t *label,
                          const float *vector, int32_t field_storage, int64_t *tissue,
                          char *msg, size_t msg_len);
/* Frees a tissue's resident fields; FPL_E_INVALID while requests use it. */
int fpl_server_remove_tissue(fpl_server *server, int64_t tissue, char *msg, size_t msg_len);
/* Validates the request (the checks of fpl_simulate) and queues it. */
int fpl_server_submit(fpl_server *server, const fpl_request *request, int64_t *ticket,
                      char *msg, size_t msg_len);
/* Validates a fiber-model request (the setup checks of fpl_simulate, the
 * storage's label width) and queues it; the generated property table is
 * checked in the slot before the march. */
int fpl_server_submit_model(fpl_server *server, const fpl_model_request *request,
                            int64_t *ticket, char *msg, size_t msg_len);
/* Waits up to timeout_s (< 0: no limit) for a ticket.  FPL_OK: the images are
 * written, warn receives the FPL_WARN_* mask (generation and march) and
 * latency_s the time from submit to completion; the ticket is then retired.
 * FPL_E_PENDING: not finished yet.  Any other code: the request failed
 * (retired). */
int fpl_server_wait(fpl_server *server, int64_t ticket, double timeout_s, uint32_t *warn,
                    double *latency_s, char *msg, size_t msg_len);
/* stats[0..8] = submitted, completed, in flight, pending, graphs captured,
 * graph replays, tissues generated, bytes the slots hold, bytes of the
 * registered tissues. */
int fpl_server_stats(fpl_server *server, int64_t stats[9]);
/* Finishes the requests in flight, fails the pending ones, frees everything. */
int fpl_server_destroy(fpl_server *server);

/* Release the library's cached device workspace of the current device. */
void fpl_release_workspace(void);

/* Library version string. */
const char *fpl_version(void);

#ifdef __cplusplus
}
#endif
#endif /* FASTPLI_AMD_H_ */
