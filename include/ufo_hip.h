/*
 * ufo_hip.h — C ABI of libufo_hip.so, the device layer under the ufo_amd
 * tracking drivers.
 *
 * It replaces the OpenCL layer of the reference: device discovery and
 * contexts (ufo/cl_utils.py:12-34), program builds (cl.Program(...).build,
 * ufo/track.py:118, ufo/sa.py:129, ufo/co.py:234, ufo/optics.py:119),
 * buffers and copies (cl.Buffer / cl.enqueue_copy, ufo/track.py:120-123,
 * 134-139) and kernel launches (kernel.run(queue, (threads,), None, ...),
 * ufo/track.py:137).  Programs are HIP C++ compiled with hipRTC for gfx950;
 * buffers are plain device pointers (the Python drivers pass torch tensor
 * data pointers); launches go to a caller-supplied HIP stream.
 *
 * Conventions: every entry point returns 0 on success or a negative
 * UFO_E_* code; the message of the last failure on the calling thread is
 * returned by ufo_last_error().  Handles are not thread-safe: use one per
 * thread or stream.
 */
#ifndef UFO_HIP_H
#define UFO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UFO_OK 0
#define UFO_E_INVALID (-1)   /* bad argument */
#define UFO_E_HIP (-2)       /* HIP runtime error */
#define UFO_E_COMPILE (-3)   /* hipRTC compile error; the log holds the compiler output */
#define UFO_E_NOTFOUND (-4)  /* no such device or kernel */
#define UFO_E_IO (-5)        /* reserved: the code-object cache is best effort (a failed
                                 * write leaves the result valid; a cached object the
                                 * runtime refuses is rebuilt) */

typedef struct ufo_ctx ufo_ctx;
typedef struct ufo_module ufo_module;
typedef struct ufo_graph ufo_graph;

typedef struct ufo_launch_cfg {
    unsigned int grid[3];
    unsigned int block[3];
    unsigned int shared_bytes;
} ufo_launch_cfg;

typedef struct ufo_device_props {
    char name[256];        /* marketing name */
    char arch[64];         /* gcnArchName, e.g. "gfx950:sramecc+:xnack-" */
    int compute_units;     /* CUs */
    int clock_khz;         /* peak shader clock */
    int warp_size;         /* 64 */
    size_t total_mem;      /* bytes of HBM */
} ufo_device_props;

/* Message of the last failing call on this thread ("" if none). */
const char *ufo_last_error(void);

/* Library version. */
int ufo_version(int *major, int *minor);

/* Number of HIP devices (replaces the platform walk of list_devices,
 * ufo/cl_utils.py:12-28). */
int ufo_device_count(int *n);

/* Name of device `dev` into buf (ufo/cl_utils.py:25, `d.name`). */
int ufo_device_name(int dev, char *buf, size_t len);

/* Compute units, clock, architecture and memory of device `dev`. */
int ufo_device_props_get(int dev, ufo_device_props *props);

/* Context bound to device `dev` (replaces cl.Context, ufo/cl_utils.py:30-34). */
int ufo_context_create(int dev, ufo_ctx **ctx);
int ufo_context_destroy(ufo_ctx *ctx);

/* Compile HIP source with hipRTC for `arch` (e.g. "gfx950") without loading
 * it: no device is needed.  On success *code receives a malloc'ed code
 * object of *size bytes (free with ufo_code_free).  The compiler log is
 * copied into log (may be NULL).  Results are cached on disk under
 * $UFO_CACHE_DIR (default ~/.cache/ufo_amd) keyed by source, options, arch
 * and hipRTC version. */
int ufo_compile(const char *src, const char *options, const char *arch,
                void **code, size_t *size, char *log, size_t loglen);
void ufo_code_free(void *code);

/* Compile (through the cache) and load a module on the context's device
 * (replaces cl.Program(ctx, src).build(options), ufo/track.py:118).
 * arch may be NULL: the device's own architecture is used. */
int ufo_module_compile(ufo_ctx *ctx, const char *src, const char *options, const char *arch,
                       ufo_module **mod, char *log, size_t loglen);

/* Load a code object produced by ufo_compile. */
int ufo_module_load(ufo_ctx *ctx, const void *code, size_t size, ufo_module **mod);
/* Unload a module.  The caller makes sure no kernel of it is still queued or
 * running (synchronise its streams first); the Python drivers keep modules
 * for the life of their context. */
int ufo_module_destroy(ufo_module *mod);

/* Register usage of a kernel: arch VGPRs, SGPRs, private bytes per lane,
 * max threads per block. */
int ufo_kernel_info(ufo_module *mod, const char *kernel, int *vgprs, int *sgprs,
                    int *private_bytes, int *max_threads);

/* Launch `kernel` with `nargs` arguments (args[i] points at the i-th
 * argument value) on `stream` (a hipStream_t of the module's device; NULL =
 * the default stream, and then the calling thread's current device must be
 * the module's, else UFO_E_INVALID); asynchronous (replaces kernel.run(queue, (threads,), None, ...),
 * ufo/track.py:137). */
int ufo_launch(ufo_module *mod, const char *kernel, const ufo_launch_cfg *cfg,
               void **args, size_t nargs, void *stream);

/* Largest scratch area ufo_launch_noalloc supplies, in bytes (1 MiB per
 * device, held from the library's load on that device until the process
 * ends). */
#define UFO_SCRATCH_MAX_BYTES (1u << 20)

/* ufo_launch for callers that cannot allocate device memory (code under
 * graph capture, serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  The kernel's scratch buffer (the work-pool
 * counter of the aperture kernel, ufo/sa.py:122, or any other zero-initialised
 * counters or staging of at most UFO_SCRATCH_MAX_BYTES) comes from a
 * device-resident array of the library instead of the caller: args[scratch_arg]
 * is ignored (may be NULL) and the kernel receives that array's device
 * address there; its first scratch_bytes (rounded up to 4) are zeroed by a
 * kernel on `stream` before the launch.  Inputs and outputs are the caller's,
 * as for ufo_launch.  Makes no hipMalloc/hipFree and no host allocation;
 * UFO_E_INVALID if scratch_bytes is 0 or above UFO_SCRATCH_MAX_BYTES, if
 * scratch_arg >= nargs or nargs > 64.  Every call on one device shares that
 * device's array, which keeps the last call's contents: calls that overlap on
 * one device (other streams, other threads) race; serialise them. */
int ufo_launch_noalloc(ufo_module *mod, const char *kernel, const ufo_launch_cfg *cfg,
                       void **args, size_t nargs, size_t scratch_arg, size_t scratch_bytes,
                       void *stream);

/* Wait for all work on `stream` (NULL = whole device). */
int ufo_sync(void *stream);

/* Device memory and copies for callers that do not bring their own
 * allocator (replaces cl.Buffer and cl.enqueue_copy,
 * ufo/track.py:120-123,134-139).  Copies are synchronous. */
int ufo_malloc(size_t bytes, void **dptr);
int ufo_free(void *dptr);
int ufo_memcpy_htod(void *dst, const void *src, size_t bytes);
int ufo_memcpy_dtoh(void *dst, const void *src, size_t bytes);

/* Stream capture into a reusable graph (turn blocks of the batched
 * tracker).  Record launches between begin and end; replay with
 * ufo_graph_launch.  No reference counterpart (the reference launches once
 * per run()). */
int ufo_graph_begin(void *stream);
int ufo_graph_end(void *stream, ufo_graph **graph);
int ufo_graph_launch(ufo_graph *graph, void *stream);
/* Free a graph.  The caller makes sure no replay of it is still queued or
 * running (synchronise the streams it was launched on first). */
int ufo_graph_destroy(ufo_graph *graph);

/* Elapsed milliseconds of `kernel` launched `reps` times back to back on
 * `stream`, measured with HIP events recorded on that stream. */
int ufo_time_launch(ufo_module *mod, const char *kernel, const ufo_launch_cfg *cfg,
                    void **args, size_t nargs, void *stream, int reps, float *ms);

#ifdef __cplusplus
}
#endif

#endif /* UFO_HIP_H */
