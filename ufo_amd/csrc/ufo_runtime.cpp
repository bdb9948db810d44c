// libufo_hip.so — the C ABI declared in include/ufo_hip.h.
//
// Device layer under the ufo_amd drivers: hipRTC builds with an on-disk
// code-object cache, module/function handles, launches on a caller's stream,
// stream-captured graphs and HIP-event timing.  Host code only; the kernels
// themselves are generated at run time by ufo_amd/codegen.py.

#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cerrno>
#include <cstdint>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include <sys/stat.h>
#include <sys/types.h>
#include <unistd.h>

#include "../../include/ufo_hip.h"

namespace {

thread_local std::string g_error;

int fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

#define HIP_TRY(call)                                                             \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess)                                                     \
            return fail(UFO_E_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

uint64_t fnv1a(uint64_t h, const void *data, size_t n) {
    const unsigned char *p = static_cast<const unsigned char *>(data);
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

std::string cache_dir() {
    const char *env = getenv("UFO_CACHE_DIR");
    if (env && *env) return env;
    const char *home = getenv("HOME");
    if (home && *home) {
        std::string d = std::string(home) + "/.cache";
        mkdir(d.c_str(), 0755);
        d += "/ufo_amd";
        if (mkdir(d.c_str(), 0755) == 0 || errno == EEXIST) {
            if (access(d.c_str(), W_OK) == 0) return d;
        }
    }
    return "/tmp/ufo_amd_cache_" + std::to_string(getuid());
}

bool read_file(const std::string &path, std::vector<char> &out) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return false;
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    if (n <= 0) {
        fclose(f);
        return false;
    }
    out.resize(static_cast<size_t>(n));
    size_t got = fread(out.data(), 1, out.size(), f);
    fclose(f);
    return got == out.size();
}

// A cached code object is handed to the loader only if its headers fit in
// the file: an ELF64 image whose program and section header tables end
// within it (a truncated entry is rebuilt instead), or an offload bundle.
bool plausible_code_object(const std::vector<char> &b) {
    static const char bundle[] = "__CLANG_OFFLOAD_BUNDLE__";
    if (b.size() >= sizeof bundle - 1 && memcmp(b.data(), bundle, sizeof bundle - 1) == 0) return true;
    if (b.size() < 64 || memcmp(b.data(), "\x7f" "ELF", 4) != 0 || b[4] != 2) return false;
    uint64_t phoff = 0, shoff = 0;
    uint16_t phentsize = 0, phnum = 0, shentsize = 0, shnum = 0;
    memcpy(&phoff, b.data() + 0x20, 8);
    memcpy(&shoff, b.data() + 0x28, 8);
    memcpy(&phentsize, b.data() + 0x36, 2);
    memcpy(&phnum, b.data() + 0x38, 2);
    memcpy(&shentsize, b.data() + 0x3A, 2);
    memcpy(&shnum, b.data() + 0x3C, 2);
    const uint64_t n = b.size();
    return phoff <= n && shoff <= n && uint64_t(phnum) * phentsize <= n - phoff &&
           uint64_t(shnum) * shentsize <= n - shoff;
}

void write_file_atomic(const std::string &path, const void *data, size_t n) {
    std::string tmp = path + ".tmp" + std::to_string(getpid());
    FILE *f = fopen(tmp.c_str(), "wb");
    if (!f) return;
    size_t put = fwrite(data, 1, n, f);
    fclose(f);
    if (put == n)
        rename(tmp.c_str(), path.c_str());
    else
        unlink(tmp.c_str());
}

void copy_log(char *log, size_t loglen, const std::string &text) {
    if (!log || loglen == 0) return;
    size_t n = text.size() < loglen - 1 ? text.size() : loglen - 1;
    memcpy(log, text.data(), n);
    log[n] = 0;
}

// Split an option string on whitespace (hipRTC takes an argv-style list).
std::vector<std::string> split_options(const char *options) {
    std::vector<std::string> out;
    if (!options) return out;
    std::string cur;
    for (const char *p = options; *p; ++p) {
        if (*p == ' ' || *p == '\t' || *p == '\n') {
            if (!cur.empty()) out.push_back(cur), cur.clear();
        } else {
            cur.push_back(*p);
        }
    }
    if (!cur.empty()) out.push_back(cur);
    return out;
}

// `cached` (optional) is set when the object came from the on-disk cache;
// `refresh` skips the cache read and overwrites the entry.
int compile_rtc(const char *src, const char *options, const std::string &arch,
                std::vector<char> &code, std::string &log, bool *cached = nullptr, bool refresh = false) {
    int ver_major = 0, ver_minor = 0;
    hiprtcVersion(&ver_major, &ver_minor);
    std::string key_text = std::string(src) + '\0' + (options ? options : "") + '\0' + arch + '\0' +
                           std::to_string(ver_major) + "." + std::to_string(ver_minor);
    uint64_t h = fnv1a(1469598103934665603ull, key_text.data(), key_text.size());
    char hex[32];
    snprintf(hex, sizeof hex, "%016llx", static_cast<unsigned long long>(h));
    std::string dir = cache_dir();
    mkdir(dir.c_str(), 0755);
    std::string path = dir + "/" + hex + ".hsaco";
    if (cached) *cached = false;
    if (!refresh && read_file(path, code) && plausible_code_object(code)) {
        if (cached) *cached = true;
        return UFO_OK;
    }

    hiprtcProgram prog;
    hiprtcResult r = hiprtcCreateProgram(&prog, src, "ufo_kernel.hip", 0, nullptr, nullptr);
    if (r != HIPRTC_SUCCESS) return fail(UFO_E_COMPILE, "hiprtcCreateProgram: %s", hiprtcGetErrorString(r));
    std::vector<std::string> opts = split_options(options);
    opts.push_back("--offload-arch=" + arch);
    std::vector<const char *> argv;
    for (auto &o : opts) argv.push_back(o.c_str());
    r = hiprtcCompileProgram(prog, static_cast<int>(argv.size()), argv.data());
    size_t log_size = 0;
    hiprtcGetProgramLogSize(prog, &log_size);
    if (log_size > 1) {
        std::vector<char> buf(log_size + 1, 0);
        hiprtcGetProgramLog(prog, buf.data());
        log.assign(buf.data());
    }
    if (r != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        return fail(UFO_E_COMPILE, "hiprtcCompileProgram: %s\n%s", hiprtcGetErrorString(r), log.c_str());
    }
    size_t size = 0;
    r = hiprtcGetCodeSize(prog, &size);
    if (r != HIPRTC_SUCCESS || size == 0) {
        hiprtcDestroyProgram(&prog);
        return fail(UFO_E_COMPILE, "hiprtcGetCodeSize: %s", hiprtcGetErrorString(r));
    }
    code.resize(size);
    r = hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    if (r != HIPRTC_SUCCESS) return fail(UFO_E_COMPILE, "hiprtcGetCode: %s", hiprtcGetErrorString(r));
    write_file_atomic(path, code.data(), code.size());
    return UFO_OK;
}

std::string base_arch(const char *gcn) {
    std::string a(gcn);
    size_t c = a.find(':');
    return c == std::string::npos ? a : a.substr(0, c);
}

}  // namespace

// Scratch memory of ufo_launch_noalloc: one array per device, allocated by
// the HIP runtime when the library's code object is loaded on that device
// and held until the process ends.  In host code `scratch` names a host-side
// stand-in, never the device memory: the device address is looked up once per
// device with hipGetSymbolAddress (scratch_address) and only that address is
// handed to kernels and fills.
static __device__ __align__(256) unsigned char scratch[UFO_SCRATCH_MAX_BYTES];

// Zero the first n 32-bit words of p (a kernel, not hipMemsetAsync, so that a
// caller may capture the call into a graph).
__global__ void ufo_scratch_clear(uint32_t *p, unsigned int n) {
    for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_lock;
void *g_scratch_ptr[kMaxDevices];   // device address of `scratch` per device, guarded by g_scratch_lock

// Device address of `scratch` on device `dev`, looked up on first use with
// `dev` current (the caller's current device is restored).
int scratch_address(int dev, void **out) {
    if (dev < 0 || dev >= kMaxDevices) return fail(UFO_E_INVALID, "device %d beyond the scratch table", dev);
    std::lock_guard<std::mutex> guard(g_scratch_lock);
    if (!g_scratch_ptr[dev]) {
        int cur = -1;
        HIP_TRY(hipGetDevice(&cur));
        if (cur != dev) HIP_TRY(hipSetDevice(dev));
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess || !ptr)
            return fail(UFO_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return UFO_OK;
}

}  // namespace

struct ufo_ctx {
    int dev;
    std::string arch;
};

struct ufo_module {
    int dev;
    hipModule_t mod;
    std::unordered_map<std::string, hipFunction_t> funcs;   // guarded by lock
    std::mutex lock;
};

struct ufo_graph {
    hipGraph_t graph;
    hipGraphExec_t exec;
};

static int lookup(ufo_module *m, const char *kernel, hipFunction_t *f) {
    std::lock_guard<std::mutex> guard(m->lock);
    auto it = m->funcs.find(kernel);
    if (it != m->funcs.end()) {
        *f = it->second;
        return UFO_OK;
    }
    hipError_t e = hipModuleGetFunction(f, m->mod, kernel);
    if (e != hipSuccess) return fail(UFO_E_NOTFOUND, "kernel '%s' not found: %s", kernel, hipGetErrorString(e));
    m->funcs[kernel] = *f;
    return UFO_OK;
}

// As lookup, without adding to the module's table (ufo_launch_noalloc makes
// no host allocation either).
static int find_function(ufo_module *m, const char *kernel, hipFunction_t *f) {
    {
        std::lock_guard<std::mutex> guard(m->lock);
        auto it = m->funcs.find(kernel);
        if (it != m->funcs.end()) {
            *f = it->second;
            return UFO_OK;
        }
    }
    hipError_t e = hipModuleGetFunction(f, m->mod, kernel);
    if (e != hipSuccess) return fail(UFO_E_NOTFOUND, "kernel '%s' not found: %s", kernel, hipGetErrorString(e));
    return UFO_OK;
}

constexpr int kMaxNoallocArgs = 64;

extern "C" {

const char *ufo_last_error(void) { return g_error.c_str(); }

int ufo_version(int *major, int *minor) {
    if (!major || !minor) return fail(UFO_E_INVALID, "null argument");
    *major = 1;
    *minor = 0;
    return UFO_OK;
}

int ufo_device_count(int *n) {
    if (!n) return fail(UFO_E_INVALID, "null argument");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e == hipErrorNoDevice) c = 0;
    else if (e != hipSuccess) return fail(UFO_E_HIP, "hipGetDeviceCount: %s", hipGetErrorString(e));
    *n = c;
    return UFO_OK;
}

int ufo_device_props_get(int dev, ufo_device_props *props) {
    if (!props) return fail(UFO_E_INVALID, "null argument");
    int n = 0;
    int rc = ufo_device_count(&n);
    if (rc) return rc;
    if (dev < 0 || dev >= n) return fail(UFO_E_NOTFOUND, "no device %d (%d present)", dev, n);
    hipDeviceProp_t p;
    HIP_TRY(hipGetDeviceProperties(&p, dev));
    memset(props, 0, sizeof *props);
    snprintf(props->name, sizeof props->name, "%s", p.name);
    snprintf(props->arch, sizeof props->arch, "%s", p.gcnArchName);
    props->compute_units = p.multiProcessorCount;
    props->clock_khz = p.clockRate;
    props->warp_size = p.warpSize;
    props->total_mem = p.totalGlobalMem;
    return UFO_OK;
}

int ufo_device_name(int dev, char *buf, size_t len) {
    if (!buf || len == 0) return fail(UFO_E_INVALID, "null buffer");
    ufo_device_props p;
    int rc = ufo_device_props_get(dev, &p);
    if (rc) return rc;
    snprintf(buf, len, "%s (%s)", p.name, p.arch);
    return UFO_OK;
}

int ufo_context_create(int dev, ufo_ctx **ctx) {
    if (!ctx) return fail(UFO_E_INVALID, "null argument");
    ufo_device_props p;
    int rc = ufo_device_props_get(dev, &p);
    if (rc) return rc;
    // set-up of the device's scratch array for ufo_launch_noalloc
    void *ptr = nullptr;
    rc = scratch_address(dev, &ptr);
    if (rc) return rc;
    *ctx = new ufo_ctx{dev, base_arch(p.arch)};
    return UFO_OK;
}

int ufo_context_destroy(ufo_ctx *ctx) {
    delete ctx;
    return UFO_OK;
}

int ufo_compile(const char *src, const char *options, const char *arch, void **code, size_t *size,
                char *log, size_t loglen) {
    if (!src || !arch || !code || !size) return fail(UFO_E_INVALID, "null argument");
    std::vector<char> obj;
    std::string text;
    int rc = compile_rtc(src, options, arch, obj, text);
    copy_log(log, loglen, text);
    if (rc) return rc;
    void *buf = malloc(obj.size());
    if (!buf) return fail(UFO_E_INVALID, "out of host memory");
    memcpy(buf, obj.data(), obj.size());
    *code = buf;
    *size = obj.size();
    return UFO_OK;
}

void ufo_code_free(void *code) { free(code); }

int ufo_module_load(ufo_ctx *ctx, const void *code, size_t size, ufo_module **mod) {
    if (!ctx || !code || !size || !mod) return fail(UFO_E_INVALID, "null argument");
    HIP_TRY(hipSetDevice(ctx->dev));
    hipModule_t m;
    HIP_TRY(hipModuleLoadData(&m, code));
    ufo_module *um = new ufo_module;
    um->dev = ctx->dev;
    um->mod = m;
    *mod = um;
    return UFO_OK;
}

int ufo_module_compile(ufo_ctx *ctx, const char *src, const char *options, const char *arch,
                       ufo_module **mod, char *log, size_t loglen) {
    if (!ctx || !src || !mod) return fail(UFO_E_INVALID, "null argument");
    std::vector<char> obj;
    std::string text;
    std::string target = arch ? std::string(arch) : ctx->arch;
    bool cached = false;
    int rc = compile_rtc(src, options, target, obj, text, &cached);
    copy_log(log, loglen, text);
    if (rc) return rc;
    rc = ufo_module_load(ctx, obj.data(), obj.size(), mod);
    if (rc == UFO_OK || !cached) return rc;
    // a cached object the runtime refuses (damaged, or written by another
    // toolchain under the same key) is rebuilt once and replaced
    rc = compile_rtc(src, options, target, obj, text, nullptr, true);
    copy_log(log, loglen, text);
    if (rc) return rc;
    return ufo_module_load(ctx, obj.data(), obj.size(), mod);
}

int ufo_module_destroy(ufo_module *mod) {
    if (!mod) return UFO_OK;
    (void)hipModuleUnload(mod->mod);
    delete mod;
    return UFO_OK;
}

int ufo_kernel_info(ufo_module *mod, const char *kernel, int *vgprs, int *sgprs, int *private_bytes,
                    int *max_threads) {
    if (!mod || !kernel) return fail(UFO_E_INVALID, "null argument");
    hipFunction_t f;
    int rc = lookup(mod, kernel, &f);
    if (rc) return rc;
    int v = 0;
    if (vgprs) {
        HIP_TRY(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_NUM_REGS, f));
        *vgprs = v;
    }
    if (sgprs) *sgprs = -1;
    if (private_bytes) {
        HIP_TRY(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, f));
        *private_bytes = v;
    }
    if (max_threads) {
        HIP_TRY(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_MAX_THREADS_PER_BLOCK, f));
        *max_threads = v;
    }
    return UFO_OK;
}

int ufo_launch(ufo_module *mod, const char *kernel, const ufo_launch_cfg *cfg, void **args, size_t nargs,
               void *stream) {
    if (!mod || !kernel || !cfg) return fail(UFO_E_INVALID, "null argument");
    if (nargs && !args) return fail(UFO_E_INVALID, "null argument list");
    for (int i = 0; i < 3; ++i)
        if (cfg->grid[i] == 0 || cfg->block[i] == 0) return fail(UFO_E_INVALID, "empty launch dimension");
    if (!stream) {
        int cur = -1;
        HIP_TRY(hipGetDevice(&cur));
        if (cur != mod->dev)
            return fail(UFO_E_INVALID, "default-stream launch on device %d of a module loaded on device %d", cur,
                        mod->dev);
    }
    hipFunction_t f;
    int rc = lookup(mod, kernel, &f);
    if (rc) return rc;
    HIP_TRY(hipModuleLaunchKernel(f, cfg->grid[0], cfg->grid[1], cfg->grid[2], cfg->block[0], cfg->block[1],
                                  cfg->block[2], cfg->shared_bytes, static_cast<hipStream_t>(stream), args,
                                  nullptr));
    return UFO_OK;
}

int ufo_launch_noalloc(ufo_module *mod, const char *kernel, const ufo_launch_cfg *cfg, void **args, size_t nargs,
                       size_t scratch_arg, size_t scratch_bytes, void *stream) {
    if (scratch_bytes == 0 || scratch_bytes > UFO_SCRATCH_MAX_BYTES)
        return fail(UFO_E_INVALID, "scratch of %zu bytes requested; the library holds 1 to %u bytes per device",
                    scratch_bytes, (unsigned)UFO_SCRATCH_MAX_BYTES);
    if (!mod || !kernel || !cfg) return fail(UFO_E_INVALID, "null argument");
    if (!args || scratch_arg >= nargs)
        return fail(UFO_E_INVALID, "scratch argument %zu outside the %zu kernel arguments", scratch_arg, nargs);
    if (nargs > kMaxNoallocArgs) return fail(UFO_E_INVALID, "%zu kernel arguments (at most %d)", nargs, kMaxNoallocArgs);
    for (int i = 0; i < 3; ++i)
        if (cfg->grid[i] == 0 || cfg->block[i] == 0) return fail(UFO_E_INVALID, "empty launch dimension");
    void *ptr = nullptr;
    int rc = scratch_address(mod->dev, &ptr);
    if (rc) return rc;
    hipFunction_t f;
    rc = find_function(mod, kernel, &f);
    if (rc) return rc;
    if (!stream) {
        int cur = -1;
        HIP_TRY(hipGetDevice(&cur));
        if (cur != mod->dev)
            return fail(UFO_E_INVALID, "default-stream launch on device %d of a module loaded on device %d", cur,
                        mod->dev);
    }
    // clear the words the kernel may use, on the launch stream, before it
    const unsigned int words = static_cast<unsigned int>((scratch_bytes + 3) / 4);
    const unsigned int blocks = (words + 255) / 256 < 64 ? (words + 255) / 256 : 64;
    hipLaunchKernelGGL(ufo_scratch_clear, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<uint32_t *>(ptr), words);
    HIP_TRY(hipGetLastError());
    // the caller's argument list (on the stack) with the scratch slot pointing at the device address
    void *local[kMaxNoallocArgs];
    for (size_t i = 0; i < nargs; ++i) local[i] = args[i];
    local[scratch_arg] = &ptr;
    HIP_TRY(hipModuleLaunchKernel(f, cfg->grid[0], cfg->grid[1], cfg->grid[2], cfg->block[0], cfg->block[1],
                                  cfg->block[2], cfg->shared_bytes, static_cast<hipStream_t>(stream), local,
                                  nullptr));
    return UFO_OK;
}

int ufo_sync(void *stream) {
    if (stream)
        HIP_TRY(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    else
        HIP_TRY(hipDeviceSynchronize());
    return UFO_OK;
}

int ufo_malloc(size_t bytes, void **dptr) {
    if (!dptr) return fail(UFO_E_INVALID, "null argument");
    HIP_TRY(hipMalloc(dptr, bytes ? bytes : 1));
    return UFO_OK;
}

int ufo_free(void *dptr) {
    if (dptr) HIP_TRY(hipFree(dptr));
    return UFO_OK;
}

int ufo_memcpy_htod(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return UFO_OK;
    if (!dst || !src) return fail(UFO_E_INVALID, "null argument");
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return UFO_OK;
}

int ufo_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return UFO_OK;
    if (!dst || !src) return fail(UFO_E_INVALID, "null argument");
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return UFO_OK;
}

int ufo_graph_begin(void *stream) {
    if (!stream) return fail(UFO_E_INVALID, "graph capture needs a non-default stream");
    HIP_TRY(hipStreamBeginCapture(static_cast<hipStream_t>(stream), hipStreamCaptureModeThreadLocal));
    return UFO_OK;
}

int ufo_graph_end(void *stream, ufo_graph **graph) {
    if (!stream || !graph) return fail(UFO_E_INVALID, "null argument");
    hipGraph_t g;
    HIP_TRY(hipStreamEndCapture(static_cast<hipStream_t>(stream), &g));
    hipGraphExec_t x;
    hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(UFO_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    *graph = new ufo_graph{g, x};
    return UFO_OK;
}

int ufo_graph_launch(ufo_graph *graph, void *stream) {
    if (!graph) return fail(UFO_E_INVALID, "null argument");
    HIP_TRY(hipGraphLaunch(graph->exec, static_cast<hipStream_t>(stream)));
    return UFO_OK;
}

int ufo_graph_destroy(ufo_graph *graph) {
    if (!graph) return UFO_OK;
    (void)hipGraphExecDestroy(graph->exec);
    (void)hipGraphDestroy(graph->graph);
    delete graph;
    return UFO_OK;
}

int ufo_time_launch(ufo_module *mod, const char *kernel, const ufo_launch_cfg *cfg, void **args, size_t nargs,
                    void *stream, int reps, float *ms) {
    if (!ms || reps < 1) return fail(UFO_E_INVALID, "bad argument");
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc = UFO_OK;
    hipError_t e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e == hipSuccess) e = hipEventRecord(a, s);
    for (int i = 0; e == hipSuccess && rc == UFO_OK && i < reps; ++i)
        rc = ufo_launch(mod, kernel, cfg, args, nargs, stream);
    if (e == hipSuccess && rc == UFO_OK) e = hipEventRecord(b, s);
    if (e == hipSuccess && rc == UFO_OK) e = hipEventSynchronize(b);
    if (e == hipSuccess && rc == UFO_OK) e = hipEventElapsedTime(ms, a, b);
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (rc) return rc;
    if (e != hipSuccess) return fail(UFO_E_HIP, "event timing: %s", hipGetErrorString(e));
    return UFO_OK;
}

}  // extern "C"
