// atsim_api.hip -- host runtime and C ABI (include/atsim.h) of the wavefront
// library: contexts, device buffers, transfer-function tables, the per-frame
// launch sequence of the layer loop.
#include <algorithm>
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <map>
#include <string>
#include <vector>

#include "atsim.h"
#include "atsim_fft.hpp"
#include "atsim_kernels.hpp"
#include "atsim_launch.hpp"

using namespace atsim;

#define ATSIM_COPY_SLOTS 4

struct atsim_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::map<long, float2*> twiddles;
    // launch-shape tuning (0 = built-in default)
    int row_mw = 0;  // wavelengths per row-pass workgroup
    int col_c = 0;   // columns per column-pass workgroup
    int col_e = 0;   // elements per thread of the column Fresnel pass (0 = default, 16, 32)
    int streams = 2; // 2..4: the Fresnel passes of that many wavelength sets run on as many streams (1: one)
    int gap_passes = 1;  // 1: one pass over the field per gap (x / y order alternating), 2: a row and a column pass
    int col2_e = 0;      // elements per thread of the column passes with a screen (0 = default, 16, 32)
    int graph = 0;   // 1: frames replay a captured HIP graph of their kernels
    // 1: with factored phase multipliers the geometric phase is sampled into
    // two planes beside the Fresnel passes and the extraction reads them;
    // 2: the extraction samples the layers itself
    int extract_planes = 1;
    // 1: the block rule's final update visits only the blocks whose largest
    // |r| can give a non-zero count; 2: every pixel of every block
    int block_apply = 1;
    // 1: frames of a simulation overlap: frame k + 1's Fresnel passes run on
    // the simulation's own stream beside frame k's extraction and block
    // fix-up, which stay on the context stream (two buffer sets alternate)
    int pipeline = 0;
    unsigned tuning_gen = 0;  // bumped by every accepted atsim_set_tuning: captured frame graphs are stale
    // bumped by every call that enqueues device writes on the context stream:
    // a pipelined frame's passes then wait for that stream first
    unsigned long work_gen = 0;
    // asynchronous device-to-host copies (default priority, like the main stream)
    hipStream_t copy_stream = nullptr;
    hipEvent_t copy_ready[ATSIM_COPY_SLOTS] = {};
    hipEvent_t copy_done[ATSIM_COPY_SLOTS] = {};
    // frames of several simulations in one call (atsim_sims_frame)
    long mg_frames = 0;
    hipEvent_t mg_fork = nullptr;
    double* nn_part = nullptr;  // 256 partial sums of the network's loss reduction
    float* nn_bias = nullptr;   // per-chunk partial sums of bias gradients
    size_t nn_bias_cap = 0;
};

// A copy ordered on the context stream and complete on return.  hipMemcpy
// alone is not ordered with the library's non-blocking streams, and a
// device-to-device (or pageable host-to-device) hipMemcpy may return before
// the data has landed, so kernels enqueued next could read it half-copied.
static hipError_t copy_on_stream(atsim_ctx* c, void* dst, const void* src, size_t nbytes) {
    hipError_t e = hipMemcpyAsync(dst, src, nbytes, hipMemcpyDefault, c->stream);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(c->stream);
}

static int fail(atsim_ctx* c, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIPCHK(ctx, expr)                                                                         \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) return fail((ctx), ATSIM_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

static bool pow2(long n) { return n > 0 && (n & (n - 1)) == 0; }

// e^{-2 pi i m / n}, m < n, built in FP64, followed (powers of two) by the
// per-stage tables of the transform plans: entries of the same table
// gathered so that consecutive k are adjacent (tw_stage_off)
static int get_twiddles(atsim_ctx* c, long n, float2** out) {
    auto it = c->twiddles.find(n);
    if (it != c->twiddles.end()) { *out = it->second; return 0; }
    const bool plans = pow2(n) && n >= 16 && n <= 8192;
    const long extra = plans ? tw_tables_size((int)n) : 0;
    std::vector<float2> h(n + extra);
    for (long m = 0; m < n; m++) {
        double a = -2.0 * M_PI * (double)m / (double)n;
        h[m] = make_float2((float)cos(a), (float)sin(a));
    }
    if (plans) {
        const int S = (ilog2c((int)n) + 3) / 4;
        for (int dir = -1; dir <= 1; dir += 2)
            for (int s = 1; s < S; s++) {
                const int R = tw_radix((int)n, s, dir), NS = tw_ns((int)n, s, dir);
                const long step = n / ((long)NS * R);
                float2* o = h.data() + n + tw_stage_off((int)n, dir, s);
                for (int e = 0; e < tw_terms(R); e++)
                    for (int k = 0; k < NS; k++) o[e * NS + k] = h[((long)1 << e) * k * step];
            }
    }
    float2* d = nullptr;
    HIPCHK(c, hipMalloc(&d, sizeof(float2) * (n + extra)));
    HIPCHK(c, copy_on_stream(c, d, h.data(), sizeof(float2) * (n + extra)));
    c->twiddles[n] = d;
    *out = d;
    return 0;
}

// 1D Fresnel transfer factor h(k) = e^{-i coeff kt^2} / n, kt = signed
// frequency index, coeff = pi z lambda / (dx n)^2 (WFpropagate.c:97-121):
// h(kx) h(ky) is the reference's quadrant-swapped 2D factor / n^2.
static void fresnel_table(float2* h, long n, double dx, double z, double lambda) {
    const double coeff = M_PI * z * lambda / (dx * n) / (dx * n);
    for (long k = 0; k < n; k++) {
        const double kt = (double)(k < n / 2 ? k : k - n);
        const double ang = -coeff * kt * kt;
        h[k] = make_float2((float)(cos(ang) / n), (float)(sin(ang) / n));
    }
}

// ------------------------------------------------------------------ transforms of any size
// In-place radix-2 FFT of a power-of-two length in FP64 on the host (the
// Bluestein filter spectrum), sign as the device transforms.
static void host_fft(std::vector<double>& re, std::vector<double>& im, int sign) {
    const size_t n = re.size();
    for (size_t i = 1, j = 0; i < n; i++) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            std::swap(re[i], re[j]);
            std::swap(im[i], im[j]);
        }
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        for (size_t i = 0; i < n; i += len)
            for (size_t k = 0; k < len / 2; k++) {
                const double a = sign * 2.0 * M_PI * (double)k / (double)len;
                const double wr = cos(a), wi = sin(a);
                const size_t p = i + k, q = i + k + len / 2;
                const double xr = re[q] * wr - im[q] * wi, xi = re[q] * wi + im[q] * wr;
                re[q] = re[p] - xr;
                im[q] = im[p] - xi;
                re[p] += xr;
                im[p] += xi;
            }
    }
}

// gather map of the reference's permut (fft.c:570-792) on an n0 x n1 image:
// after the swap loops, element p holds what was at sigma[p]
static std::vector<int> permut_sigma(long n0, long n1) {
    std::vector<int> idx((size_t)(n0 * n1));
    for (long k = 0; k < n0 * n1; k++) idx[(size_t)k] = (int)k;
    const long xh = n0 / 2, yh = n1 / 2;
    for (long jj = 0; jj < yh; jj++)
        for (long ii = 0; ii < xh; ii++) std::swap(idx[(size_t)(jj * n0 + ii)], idx[(size_t)((jj + yh) * n0 + ii + xh)]);
    for (long jj = yh; jj < n1; jj++)
        for (long ii = 0; ii < xh; ii++) std::swap(idx[(size_t)(jj * n0 + ii)], idx[(size_t)((jj - yh) * n0 + ii + xh)]);
    return idx;
}

static long next_pow2(long n) {
    long m = 1;
    while (m < n) m <<= 1;
    return m;
}

// ------------------------------------------------------------------ allocation-free Fresnel
// Device memory of atsim_fresnel_noalloc: the 1D transfer factor h(k) and the
// twiddles e^{-2 pi i m / n} of one size, n <= ATSIM_NOALLOC_MAX_N (128 KB per
// device, held until the process ends).  Every call on a device shares its
// copy, and calls that overlap on one device race on it.
#define ATSIM_NOALLOC_MAX_N 8192
#define ATSIM_MAX_DEVICES 64
static __device__ __align__(256) float2 scratch[2 * ATSIM_NOALLOC_MAX_N];
// device address of `scratch` per device, from hipGetSymbolAddress at the
// first context on that device (the host-side name is only a stand-in)
static float2* g_scratch[ATSIM_MAX_DEVICES];

// tab[k] = h(k) as fresnel_table builds it, tw[m] = e^{-2 pi i m / n} as
// get_twiddles builds it, both in FP64 and rounded to float
__global__ void k_fresnel_tables(float2* __restrict__ tab, float2* __restrict__ tw, long n, double coeff) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double kt = (double)(k < n / 2 ? k : k - n);
    const double ang = -coeff * kt * kt;
    tab[k] = make_float2((float)(cos(ang) / n), (float)(sin(ang) / n));
    const double a = -2.0 * M_PI * (double)k / (double)n;
    tw[k] = make_float2((float)cos(a), (float)sin(a));
}

extern "C" {

int atsim_ctx_create(int device, atsim_ctx** out) {
    if (!out) return ATSIM_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return ATSIM_EHIP;
    if (device < 0 || device >= ndev) return ATSIM_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return ATSIM_EHIP;
    if (device < ATSIM_MAX_DEVICES && !g_scratch[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return ATSIM_EHIP;
        g_scratch[device] = (float2*)p;
    }
    atsim_ctx* c = new atsim_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return ATSIM_EHIP;
    }
    *out = c;
    return ATSIM_OK;
}

void atsim_ctx_destroy(atsim_ctx* c) {
    if (!c) return;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    for (auto& kv : c->twiddles) hipFree(kv.second);
    if (c->mg_fork) hipEventDestroy(c->mg_fork);
    if (c->nn_part) hipFree(c->nn_part);
    if (c->nn_bias) hipFree(c->nn_bias);
    if (c->copy_stream) {
        hipStreamSynchronize(c->copy_stream);
        for (int i = 0; i < ATSIM_COPY_SLOTS; i++) {
            hipEventDestroy(c->copy_ready[i]);
            hipEventDestroy(c->copy_done[i]);
        }
        hipStreamDestroy(c->copy_stream);
    }
    hipStreamDestroy(c->stream);
    delete c;
}

const char* atsim_error(const atsim_ctx* c) { return c ? c->err.c_str() : "null context"; }

int atsim_sync(atsim_ctx* c) {
    if (!c) return ATSIM_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ATSIM_OK;
}

void* atsim_stream(atsim_ctx* c) { return c ? (void*)c->stream : nullptr; }

int atsim_malloc(atsim_ctx* c, void** ptr, size_t nbytes) {
    if (!c || !ptr) return ATSIM_EINVAL;
    *ptr = nullptr;
    HIPCHK(c, hipSetDevice(c->device));
    if (hipMalloc(ptr, nbytes > 0 ? nbytes : 1) != hipSuccess) {
        (void)hipGetLastError();
        return fail(c, ATSIM_ENOMEM, "cannot allocate %zu device bytes", nbytes);
    }
    return ATSIM_OK;
}

int atsim_free(atsim_ctx* c, void* ptr) {
    if (!c) return ATSIM_EINVAL;
    if (ptr) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipFree(ptr));
    }
    return ATSIM_OK;
}

int atsim_copy(atsim_ctx* c, void* dst, const void* src, size_t nbytes) {
    if (!c || !dst || !src) return ATSIM_EINVAL;
    c->work_gen++;
    HIPCHK(c, copy_on_stream(c, dst, src, nbytes));
    return ATSIM_OK;
}

int atsim_host_alloc(atsim_ctx* c, void** ptr, size_t nbytes) {
    if (!c || !ptr) return ATSIM_EINVAL;
    *ptr = nullptr;
    if (hipHostMalloc(ptr, nbytes > 0 ? nbytes : 1, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        return fail(c, ATSIM_ENOMEM, "cannot allocate %zu pinned host bytes", nbytes);
    }
    return ATSIM_OK;
}

int atsim_host_free(atsim_ctx* c, void* ptr) {
    if (!c) return ATSIM_EINVAL;
    if (ptr) HIPCHK(c, hipHostFree(ptr));
    return ATSIM_OK;
}

static int copy_stream_init(atsim_ctx* c) {
    if (c->copy_stream) return ATSIM_OK;
    HIPCHK(c, hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
    for (int i = 0; i < ATSIM_COPY_SLOTS; i++) {
        HIPCHK(c, hipEventCreateWithFlags(&c->copy_ready[i], hipEventDisableTiming));
        HIPCHK(c, hipEventCreateWithFlags(&c->copy_done[i], hipEventDisableTiming));
    }
    return ATSIM_OK;
}

int atsim_copy_async(atsim_ctx* c, int slot, void* dst, const void* src, size_t nbytes) {
    if (!c || !dst || !src || slot < 0 || slot >= ATSIM_COPY_SLOTS) return ATSIM_EINVAL;
    if (int r = copy_stream_init(c)) return r;
    HIPCHK(c, hipEventRecord(c->copy_ready[slot], c->stream));
    HIPCHK(c, hipStreamWaitEvent(c->copy_stream, c->copy_ready[slot], 0));
    HIPCHK(c, hipMemcpyAsync(dst, src, nbytes, hipMemcpyDefault, c->copy_stream));
    HIPCHK(c, hipEventRecord(c->copy_done[slot], c->copy_stream));
    return ATSIM_OK;
}

int atsim_copy_wait(atsim_ctx* c, int slot) {
    if (!c || slot < 0 || slot >= ATSIM_COPY_SLOTS) return ATSIM_EINVAL;
    if (!c->copy_stream) return ATSIM_OK;
    HIPCHK(c, hipEventSynchronize(c->copy_done[slot]));
    return ATSIM_OK;
}

int atsim_stream_wait_copy(atsim_ctx* c, int slot) {
    if (!c || slot < 0 || slot >= ATSIM_COPY_SLOTS) return ATSIM_EINVAL;
    if (!c->copy_stream) return ATSIM_OK;
    HIPCHK(c, hipStreamWaitEvent(c->stream, c->copy_done[slot], 0));
    return ATSIM_OK;
}

int atsim_set_tuning(atsim_ctx* c, const char* key, int value) {
    if (!c || !key) return ATSIM_EINVAL;
    if (!strcmp(key, "row_mw")) {
        if (value < 0 || value > 8) return fail(c, ATSIM_EINVAL, "row_mw must be in [0, 8]");
        c->row_mw = value;
    } else if (!strcmp(key, "col_c")) {
        if (value != 0 && value != 2 && value != 4 && value != 8 && value != 16)
            return fail(c, ATSIM_EINVAL, "col_c must be 0, 2, 4, 8 or 16");
        c->col_c = value;
    } else if (!strcmp(key, "col_e")) {
        if (value != 0 && value != 16 && value != 32) return fail(c, ATSIM_EINVAL, "col_e must be 0, 16 or 32");
        c->col_e = value;
    } else if (!strcmp(key, "streams")) {
        if (value < 0 || value > 4) return fail(c, ATSIM_EINVAL, "streams must be in [0, 4]");
        c->streams = value == 0 ? 2 : value;
    } else if (!strcmp(key, "col2_e")) {
        if (value != 0 && value != 16 && value != 32) return fail(c, ATSIM_EINVAL, "col2_e must be 0, 16 or 32");
        c->col2_e = value;
    } else if (!strcmp(key, "gap_passes")) {
        if (value < 0 || value > 2) return fail(c, ATSIM_EINVAL, "gap_passes must be 0, 1 or 2");
        c->gap_passes = value == 0 ? 1 : value;
    } else if (!strcmp(key, "graph")) {
        if (value != 0 && value != 1) return fail(c, ATSIM_EINVAL, "graph must be 0 or 1");
        c->graph = value;
    } else if (!strcmp(key, "block_apply")) {
        if (value < 0 || value > 2) return fail(c, ATSIM_EINVAL, "block_apply must be 0, 1 or 2");
        c->block_apply = value == 0 ? 1 : value;
    } else if (!strcmp(key, "extract_planes")) {
        if (value < 0 || value > 2) return fail(c, ATSIM_EINVAL, "extract_planes must be 0, 1 or 2");
        c->extract_planes = value == 0 ? 1 : value;
    } else if (!strcmp(key, "pipeline")) {
        if (value != 0 && value != 1) return fail(c, ATSIM_EINVAL, "pipeline must be 0 or 1");
        c->pipeline = value;
    } else {
        return fail(c, ATSIM_EINVAL, "unknown tuning key '%s'", key);
    }
    c->tuning_gen++;
    return ATSIM_OK;
}

int atsim_get_tuning(atsim_ctx* c, const char* key) {
    if (!c || !key) return ATSIM_EINVAL;
    if (!strcmp(key, "row_mw")) return c->row_mw;
    if (!strcmp(key, "col_c")) return c->col_c;
    if (!strcmp(key, "col_e")) return c->col_e;
    if (!strcmp(key, "streams")) return c->streams;
    if (!strcmp(key, "gap_passes")) return c->gap_passes;
    if (!strcmp(key, "col2_e")) return c->col2_e;
    if (!strcmp(key, "graph")) return c->graph;
    if (!strcmp(key, "extract_planes")) return c->extract_planes;
    if (!strcmp(key, "block_apply")) return c->block_apply;
    if (!strcmp(key, "pipeline")) return c->pipeline;
    return ATSIM_EINVAL;
}

int atsim_device_ok(atsim_ctx* c) {
    if (!c) return 0;
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, c->device) != hipSuccess) return 0;
    return strncmp(p.gcnArchName, "gfx950", 6) == 0 ? 1 : 0;
}

// ------------------------------------------------------------------ FFT
int atsim_fft2d(atsim_ctx* c, void* d_data, long n, long nbatch, int sign) {
    if (!c || !d_data) return ATSIM_EINVAL;
    c->work_gen++;
    if (!supported_n(n)) return fail(c, ATSIM_EUNSUPPORTED, "fft2d: size %ld is not a power of two in [16, 8192]", n);
    if (nbatch <= 0 || nbatch > 65535) return fail(c, ATSIM_EINVAL, "fft2d: nbatch %ld out of range", nbatch);
    if (sign != -1 && sign != 1) return fail(c, ATSIM_EINVAL, "fft2d: sign must be -1 or +1");
    float2* tw;
    if (int r = get_twiddles(c, n, &tw)) return r;
    ScreenArgs none{};
    HIPCHK(c, launch_row(n, OP_FFT, sign, false, (float2*)d_data, nbatch, nullptr, tw, none, c->stream));
    HIPCHK(c, launch_col(n, OP_FFT, sign, (float2*)d_data, nbatch, nullptr, 1, tw, c->stream));
    return ATSIM_OK;
}

int atsim_fft2d_variant(atsim_ctx* c, void* d_data, long n, long nbatch, int sign, int variant) {
    if (!c || !d_data) return ATSIM_EINVAL;
    c->work_gen++;
    if (variant == 0) return atsim_fft2d(c, d_data, n, nbatch, sign);
    if (variant != 1) return fail(c, ATSIM_EINVAL, "fft2d_variant: unknown variant %d", variant);
    if (!fft_mfma_supported(n))
        return fail(c, ATSIM_EUNSUPPORTED, "fft2d_variant: the matrix-core variant is built for n = 2048 (got %ld)", n);
    if (nbatch <= 0 || nbatch > 65535) return fail(c, ATSIM_EINVAL, "fft2d_variant: nbatch %ld out of range", nbatch);
    if (sign != -1 && sign != 1) return fail(c, ATSIM_EINVAL, "fft2d_variant: sign must be -1 or +1");
    float2* tw;
    if (int r = get_twiddles(c, n, &tw)) return r;
    float2* d = (float2*)d_data;
    float2* tmp = nullptr;
    HIPCHK(c, hipMallocAsync((void**)&tmp, sizeof(float2) * n * n * nbatch, c->stream));
    int r = ATSIM_OK;
    // rows, transpose, rows (the columns), transpose back
    if (launch_row_fft_mfma(d, n * nbatch, sign, tw, c->stream) != hipSuccess ||
        launch_transpose(d, tmp, n, nbatch, c->stream) != hipSuccess ||
        launch_row_fft_mfma(tmp, n * nbatch, sign, tw, c->stream) != hipSuccess ||
        launch_transpose(tmp, d, n, nbatch, c->stream) != hipSuccess)
        r = fail(c, ATSIM_EHIP, "fft2d_variant: launch failed");
    hipFreeAsync(tmp, c->stream);
    return r;
}

int atsim_row_fft_variant(atsim_ctx* c, void* d_data, long n, long nrows, int sign, int variant) {
    if (!c || !d_data) return ATSIM_EINVAL;
    c->work_gen++;
    if (sign != -1 && sign != 1) return fail(c, ATSIM_EINVAL, "row_fft_variant: sign must be -1 or +1");
    if (nrows <= 0 || nrows % n) return fail(c, ATSIM_EINVAL, "row_fft_variant: nrows must be a multiple of n");
    float2* tw;
    if (int r = get_twiddles(c, n, &tw)) return r;
    if (variant == 0) {
        if (!supported_n(n)) return fail(c, ATSIM_EUNSUPPORTED, "row_fft_variant: size %ld unsupported", n);
        ScreenArgs none{};
        HIPCHK(c, launch_row(n, OP_FFT, sign, false, (float2*)d_data, nrows / n, nullptr, tw, none, c->stream));
        return ATSIM_OK;
    }
    if (variant != 1 || !fft_mfma_supported(n))
        return fail(c, ATSIM_EUNSUPPORTED, "row_fft_variant: variant %d at n = %ld is not built", variant, n);
    HIPCHK(c, launch_row_fft_mfma((float2*)d_data, nrows, sign, tw, c->stream));
    return ATSIM_OK;
}

int atsim_permut(atsim_ctx* c, void* d_data, long n0, long n1, long nbatch) {
    if (!c || !d_data) return ATSIM_EINVAL;
    c->work_gen++;
    if (n0 <= 0 || n1 <= 0 || nbatch <= 0) return fail(c, ATSIM_EINVAL, "permut: sizes must be > 0");
    if (!(n0 & 1) && !(n1 & 1)) {
        HIPCHK(c, launch_permut((float2*)d_data, n0, n1, nbatch, c->stream));
        return ATSIM_OK;
    }
    // odd sizes: the swap loops are no longer a half-size shift; their
    // gather map is built on the host and applied on the device
    if (n0 * n1 > (1L << 30)) return fail(c, ATSIM_EUNSUPPORTED, "permut: image too large");
    const std::vector<int> sig = permut_sigma(n0, n1);
    int* d_sig = nullptr;
    float2* tmp = nullptr;
    const size_t bytes = sizeof(float2) * n0 * n1 * nbatch;
    HIPCHK(c, hipMallocAsync((void**)&d_sig, sizeof(int) * sig.size(), c->stream));
    int r = ATSIM_OK;
    if (hipMallocAsync((void**)&tmp, bytes, c->stream) != hipSuccess) r = fail(c, ATSIM_ENOMEM, "permut: no memory");
    if (r == ATSIM_OK && (hipMemcpyAsync(d_sig, sig.data(), sizeof(int) * sig.size(), hipMemcpyHostToDevice, c->stream) !=
                              hipSuccess ||
                          hipMemcpyAsync(tmp, d_data, bytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
                          launch_index_gather(tmp, (float2*)d_data, d_sig, n0 * n1, nbatch, c->stream) != hipSuccess))
        r = fail(c, ATSIM_EHIP, "permut: device gather failed");
    if (tmp) hipFreeAsync(tmp, c->stream);
    hipFreeAsync(d_sig, c->stream);
    HIPCHK(c, hipStreamSynchronize(c->stream));  // the host map must outlive its copy
    return r;
}

// One 1D transform of length L (unnormalised, sign) over nlines lines of a
// batch of images (line base (l / per_img) img_size + (l % per_img) lstep,
// elements estride apart), in place.  Power-of-two L: gather into padded
// rows, row FFT, scatter.  Other L: Bluestein -- with c_n = e^{s i pi n^2 / L}
// and d_m = e^{-s i pi m^2 / L}, X_k = c_k sum_n (x_n c_n) d_{k-n}, the
// convolution taken by FFTs of length M >= 2L - 1 (filter spectrum built on
// the host in FP64, 1/M folded in).  Chirp angles use n^2 mod 2L (exact).
static int axis_transform(atsim_ctx* c, float2* data, long L, long nlines, long per_img, long img_size, long lstep,
                          long estride, int sign) {
    if (L == 1) return ATSIM_OK;
    const bool p2 = pow2(L);
    const long M = p2 ? L : next_pow2(2 * L - 1);
    // the device row transforms start at 16: shorter lines take the padded
    // (Bluestein) path at length 16 >= 2 L - 1
    const long Mf = std::max<long>(M, 16);
    if (Mf > 8192)
        return fail(c, ATSIM_EUNSUPPORTED, "fft: axis length %ld needs a transform of %ld > 8192", L, Mf);
    const bool blue = !p2 || M < 16;
    const long F = (nlines + Mf - 1) / Mf;
    if (F > 65535) return fail(c, ATSIM_EUNSUPPORTED, "fft: %ld lines of length %ld are too many", nlines, L);
    float2* tw = nullptr;
    if (int r = get_twiddles(c, Mf, &tw)) return r;
    float2 *scr = nullptr, *d_c = nullptr, *d_h = nullptr;
    HIPCHK(c, hipMallocAsync((void**)&scr, sizeof(float2) * F * Mf * Mf, c->stream));
    std::vector<float2> hc, hh;
    int r = ATSIM_OK;
    if (blue) {
        hc.resize((size_t)L);
        std::vector<double> dr((size_t)Mf, 0.0), di((size_t)Mf, 0.0);
        for (long n = 0; n < L; n++) {
            const long q = (long)(((unsigned long long)n * (unsigned long long)n) % (unsigned long long)(2 * L));
            const double a = M_PI * (double)q / (double)L;
            hc[(size_t)n] = make_float2((float)cos(a), (float)(sign * sin(a)));
            dr[(size_t)n] = cos(a);
            di[(size_t)n] = -sign * sin(a);
            if (n > 0) {
                dr[(size_t)(Mf - n)] = cos(a);
                di[(size_t)(Mf - n)] = -sign * sin(a);
            }
        }
        host_fft(dr, di, -1);
        hh.resize((size_t)Mf);
        for (long k = 0; k < Mf; k++) hh[(size_t)k] = make_float2((float)(dr[(size_t)k] / Mf), (float)(di[(size_t)k] / Mf));
        if (hipMallocAsync((void**)&d_c, sizeof(float2) * L, c->stream) != hipSuccess ||
            hipMallocAsync((void**)&d_h, sizeof(float2) * Mf, c->stream) != hipSuccess ||
            hipMemcpyAsync(d_c, hc.data(), sizeof(float2) * L, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
            hipMemcpyAsync(d_h, hh.data(), sizeof(float2) * Mf, hipMemcpyHostToDevice, c->stream) != hipSuccess)
            r = fail(c, ATSIM_EHIP, "fft: table upload failed");
    }
    ScreenArgs none{};
    auto ok = [](hipError_t e) { return e == hipSuccess; };
    if (r == ATSIM_OK) {
        bool good = ok(launch_line_gather(data, scr, nlines, L, Mf, per_img, img_size, lstep, estride, d_c, c->stream));
        if (blue) {
            good = good && ok(launch_row(Mf, OP_FFT, -1, false, scr, F, nullptr, tw, none, c->stream)) &&
                   ok(launch_row_mul(scr, d_h, F * Mf, Mf, c->stream)) &&
                   ok(launch_row(Mf, OP_FFT, +1, false, scr, F, nullptr, tw, none, c->stream));
        } else {
            good = good && ok(launch_row(Mf, OP_FFT, sign, false, scr, F, nullptr, tw, none, c->stream));
        }
        good = good && ok(launch_line_scatter(scr, data, nlines, L, Mf, per_img, img_size, lstep, estride, d_c,
                                              c->stream));
        if (!good) r = fail(c, ATSIM_EHIP, "fft: launch failed for length %ld", L);
    }
    hipFreeAsync(scr, c->stream);
    if (d_c) hipFreeAsync(d_c, c->stream);
    if (d_h) hipFreeAsync(d_h, c->stream);
    HIPCHK(c, hipStreamSynchronize(c->stream));  // host tables must outlive their copies
    return r;
}

int atsim_fft2d_nm(atsim_ctx* c, void* d_data, long n0, long n1, long nbatch, int sign) {
    if (!c || !d_data) return ATSIM_EINVAL;
    c->work_gen++;
    if (n0 <= 0 || n1 <= 0 || nbatch <= 0) return fail(c, ATSIM_EINVAL, "fft2d_nm: sizes must be > 0");
    if (sign != -1 && sign != 1) return fail(c, ATSIM_EINVAL, "fft2d_nm: sign must be -1 or +1");
    if (n0 == n1 && supported_n(n0) && nbatch <= 65535) return atsim_fft2d(c, d_data, n0, nbatch, sign);
    float2* d = (float2*)d_data;
    // along x: lines are the rows, n1 per image
    if (int r = axis_transform(c, d, n0, n1 * nbatch, n1, n0 * n1, n0, 1, sign)) return r;
    // along y: lines are the columns, n0 per image
    return axis_transform(c, d, n1, n0 * nbatch, n0, n0 * n1, 1, n0, sign);
}

// ------------------------------------------------------------------ Fresnel
static int fresnel_run(atsim_ctx* c, float2* d, long n, long nbatch, const float2* d_tab) {
    float2* tw;
    if (int r = get_twiddles(c, n, &tw)) return r;
    ScreenArgs none{};
    none.W = 1;
    HIPCHK(c, launch_row(n, OP_FRESNEL, 0, false, d, nbatch, d_tab, tw, none, c->stream));
    HIPCHK(c, launch_col(n, OP_FRESNEL, 0, d, nbatch, d_tab, 1, tw, c->stream));
    return ATSIM_OK;
}

int atsim_fresnel(atsim_ctx* c, const void* d_in, void* d_out, long n, long nbatch, double dx, double z,
                  double lambda) {
    if (!c || !d_in || !d_out) return ATSIM_EINVAL;
    c->work_gen++;
    if (!supported_n(n)) return fail(c, ATSIM_EUNSUPPORTED, "fresnel: size %ld is not a power of two in [16, 8192]", n);
    if (nbatch <= 0 || nbatch > 65535) return fail(c, ATSIM_EINVAL, "fresnel: nbatch %ld out of range", nbatch);
    if (!(dx > 0.0)) return fail(c, ATSIM_EINVAL, "fresnel: pupil scale must be > 0");
    std::vector<float2> h(n);
    fresnel_table(h.data(), n, dx, z, lambda);
    float2* d_tab = nullptr;
    HIPCHK(c, hipMallocAsync((void**)&d_tab, sizeof(float2) * n, c->stream));
    HIPCHK(c, hipMemcpyAsync(d_tab, h.data(), sizeof(float2) * n, hipMemcpyHostToDevice, c->stream));
    if (d_in != d_out)
        HIPCHK(c, hipMemcpyAsync(d_out, d_in, sizeof(float2) * n * n * nbatch, hipMemcpyDeviceToDevice, c->stream));
    int r = fresnel_run(c, (float2*)d_out, n, nbatch, d_tab);
    hipFreeAsync(d_tab, c->stream);
    HIPCHK(c, hipStreamSynchronize(c->stream));  // host table must outlive the copy
    return r;
}

int atsim_fresnel_nm(atsim_ctx* c, const void* d_in, void* d_out, long n0, long n1, long nbatch, double dx,
                     double z, double lambda) {
    if (!c || !d_in || !d_out) return ATSIM_EINVAL;
    c->work_gen++;
    if (n0 <= 0 || n1 <= 0 || nbatch <= 0) return fail(c, ATSIM_EINVAL, "fresnel_nm: sizes must be > 0");
    if (!(dx > 0.0)) return fail(c, ATSIM_EINVAL, "fresnel_nm: pupil scale must be > 0");
    if (n0 == n1 && supported_n(n0) && nbatch <= 65535) return atsim_fresnel(c, d_in, d_out, n0, nbatch, dx, z, lambda);
    const size_t bytes = sizeof(float2) * n0 * n1 * nbatch;
    float2* tmp = nullptr;
    int* d_sig = nullptr;
    std::vector<int> sig;
    HIPCHK(c, hipMallocAsync((void**)&tmp, bytes, c->stream));
    int r = ATSIM_OK;
    if (hipMemcpyAsync(tmp, d_in, bytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
        r = fail(c, ATSIM_EHIP, "fresnel_nm: copy failed");
    if (r == ATSIM_OK) r = atsim_fft2d_nm(c, tmp, n0, n1, nbatch, -1);
    if (r == ATSIM_OK && ((n0 & 1) || (n1 & 1))) {
        sig = permut_sigma(n0, n1);
        if (hipMallocAsync((void**)&d_sig, sizeof(int) * sig.size(), c->stream) != hipSuccess ||
            hipMemcpyAsync(d_sig, sig.data(), sizeof(int) * sig.size(), hipMemcpyHostToDevice, c->stream) != hipSuccess)
            r = fail(c, ATSIM_EHIP, "fresnel_nm: permut map upload failed");
    }
    // WFpropagate.c:97: the coefficient uses naxes[0] for both axes
    const double coeff = M_PI * z * lambda / (dx * n0) / (dx * n0);
    if (r == ATSIM_OK && launch_fresnel_any(tmp, (float2*)d_out, n0, n1, nbatch, coeff, d_sig, c->stream) != hipSuccess)
        r = fail(c, ATSIM_EHIP, "fresnel_nm: launch failed");
    if (r == ATSIM_OK) r = atsim_fft2d_nm(c, d_out, n0, n1, nbatch, +1);
    hipFreeAsync(tmp, c->stream);
    if (d_sig) hipFreeAsync(d_sig, c->stream);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return r;
}

int atsim_fresnel_noalloc(atsim_ctx* c, const void* d_in, void* d_out, long n, long nbatch, double dx, double z,
                          double lambda) {
    if (!c || !d_in || !d_out) return ATSIM_EINVAL;
    c->work_gen++;
    if (n > ATSIM_NOALLOC_MAX_N)
        return fail(c, ATSIM_EUNSUPPORTED, "fresnel_noalloc: size %ld exceeds the maximum %d", n, ATSIM_NOALLOC_MAX_N);
    if (!supported_n(n))
        return fail(c, ATSIM_EUNSUPPORTED, "fresnel_noalloc: size %ld is not a power of two in [16, 8192]", n);
    if (nbatch <= 0 || nbatch > 65535) return fail(c, ATSIM_EINVAL, "fresnel_noalloc: nbatch %ld out of range", nbatch);
    if (!(dx > 0.0)) return fail(c, ATSIM_EINVAL, "fresnel_noalloc: pupil scale must be > 0");
    float2* base = c->device < ATSIM_MAX_DEVICES ? g_scratch[c->device] : nullptr;
    if (!base) return fail(c, ATSIM_EUNSUPPORTED, "fresnel_noalloc: no scratch memory on device %d", c->device);
    float2* d_tab = base;
    float2* d_tw = base + ATSIM_NOALLOC_MAX_N;
    HIPCHK(c, hipSetDevice(c->device));
    const double coeff = M_PI * z * lambda / (dx * n) / (dx * n);
    k_fresnel_tables<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(d_tab, d_tw, n, coeff);
    HIPCHK(c, hipGetLastError());
    if (d_in != d_out)
        HIPCHK(c, hipMemcpyAsync(d_out, d_in, sizeof(float2) * n * n * nbatch, hipMemcpyDeviceToDevice, c->stream));
    ScreenArgs none{};
    none.W = 1;
    HIPCHK(c, launch_row(n, OP_FRESNEL, 0, false, (float2*)d_out, nbatch, d_tab, d_tw, none, c->stream));
    HIPCHK(c, launch_col(n, OP_FRESNEL, 0, (float2*)d_out, nbatch, d_tab, 1, d_tw, c->stream));
    return ATSIM_OK;
}

int atsim_fresnel_tf_init(atsim_ctx* c, void* d_C, long n, double dx, double z, double lambda, double fpmaskrad) {
    if (!c || !d_C) return ATSIM_EINVAL;
    c->work_gen++;
    if (n <= 0 || (n & 1)) return fail(c, ATSIM_EUNSUPPORTED, "fresnel_tf_init: even size required (got %ld)", n);
    const double coeff = M_PI * z * lambda / (dx * n) / (dx * n);
    HIPCHK(c, launch_tf_init((float2*)d_C, n, coeff, fpmaskrad, c->stream));
    return ATSIM_OK;
}

int atsim_fresnel_tf_apply(atsim_ctx* c, const void* d_in, void* d_out, const void* d_C, long n, long nbatch) {
    if (!c || !d_in || !d_out || !d_C) return ATSIM_EINVAL;
    c->work_gen++;
    if (!supported_n(n)) return fail(c, ATSIM_EUNSUPPORTED, "fresnel_tf_apply: size %ld unsupported", n);
    if (d_in != d_out)
        HIPCHK(c, hipMemcpyAsync(d_out, d_in, sizeof(float2) * n * n * nbatch, hipMemcpyDeviceToDevice, c->stream));
    if (int r = atsim_fft2d(c, d_out, n, nbatch, -1)) return r;
    HIPCHK(c, launch_cmul((float2*)d_out, (const float2*)d_C, n * n, nbatch, c->stream));
    return atsim_fft2d(c, d_out, n, nbatch, +1);
}

int atsim_fresnel_cube(atsim_ctx* c, const void* d_in, float* d_amp, float* d_pha, long n, double dx,
                       double zstart, double zend, long nz, double lambda) {
    if (!c || !d_in || !d_amp || !d_pha) return ATSIM_EINVAL;
    c->work_gen++;
    if (!supported_n(n)) return fail(c, ATSIM_EUNSUPPORTED, "fresnel_cube: size %ld unsupported", n);
    if (nz <= 0) return fail(c, ATSIM_EINVAL, "fresnel_cube: nz must be > 0");
    float2* tmp = nullptr;
    HIPCHK(c, hipMalloc(&tmp, sizeof(float2) * n * n));
    int r = ATSIM_OK;
    for (long k = 0; k < nz && r == ATSIM_OK; k++) {
        const double zp = zstart + (zend - zstart) * k / nz;
        r = atsim_fresnel(c, d_in, tmp, n, 1, dx, zp, lambda);
        if (r == ATSIM_OK) {
            hipError_t e = launch_amp_pha(tmp, d_amp + k * n * n, d_pha + k * n * n, n * n, c->stream);
            if (e != hipSuccess) r = fail(c, ATSIM_EHIP, "amp_pha: %s", hipGetErrorString(e));
        }
    }
    hipStreamSynchronize(c->stream);
    hipFree(tmp);
    return r;
}

// ------------------------------------------------------------------ noise
int atsim_noise_c64(atsim_ctx* c, void* d_out, long n, uint64_t seed, uint32_t stream) {
    if (!c || !d_out || n <= 0) return ATSIM_EINVAL;
    c->work_gen++;
    HIPCHK(c, launch_noise((float2*)d_out, n, seed, stream, c->stream));
    return ATSIM_OK;
}

// ------------------------------------------------------------------ screens
static int reduce_partials(atsim_ctx* c, double* d_part, int nblocks, int nv, std::vector<double>& out) {
    std::vector<double> h((size_t)nblocks * nv);
    HIPCHK(c, hipMemcpyAsync(h.data(), d_part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    out.assign(nv, 0.0);
    for (int b = 0; b < nblocks; b++)
        for (int k = 0; k < nv; k++) out[k] += h[(size_t)b * nv + k];
    return 0;
}

int atsim_make_master(atsim_ctx* c, const void* d_noise, long size, double outerscale_px, double innerscale_px,
                      double rlim, float* d_out1, float* d_out2, double* C1, double* C2) {
    if (!c || !d_noise || !d_out1) return ATSIM_EINVAL;
    c->work_gen++;
    if (!supported_n(size)) return fail(c, ATSIM_EUNSUPPORTED, "make_master: size %ld unsupported", size);
    if (!(outerscale_px > 0) || !(innerscale_px > 0)) return fail(c, ATSIM_EINVAL, "make_master: scales must be > 0");
    const double f0 = 1.0 * size / outerscale_px;                        // AT.c:679
    const double fi = (5.92 / (2.0 * M_PI)) * size / innerscale_px;       // AT.c:680
    float2* buf = nullptr;
    const int nblocks = 1024;
    double* part = nullptr;
    HIPCHK(c, hipMalloc(&buf, sizeof(float2) * size * size));
    if (hipMalloc(&part, sizeof(double) * nblocks * 8) != hipSuccess) {
        hipFree(buf);
        return fail(c, ATSIM_ENOMEM, "make_master: allocation failed");
    }
    int r = ATSIM_OK;
    std::vector<double> s;
    hipError_t e = launch_synth_filter((const float2*)d_noise, buf, size, f0, fi, rlim, c->stream);
    if (e != hipSuccess) r = fail(c, ATSIM_EHIP, "synth_filter: %s", hipGetErrorString(e));
    if (r == ATSIM_OK) r = atsim_fft2d(c, buf, size, 1, -1);
    if (r == ATSIM_OK) {
        e = launch_lag_sums(buf, size, part, nblocks, c->stream);
        if (e != hipSuccess) r = fail(c, ATSIM_EHIP, "lag_sums: %s", hipGetErrorString(e));
    }
    if (r == ATSIM_OK) r = reduce_partials(c, part, nblocks, 8, s);
    if (r == ATSIM_OK) {
        // AT.c:794-847: lags (ii, jj) in {1,2}^2, ii outer; sums ordered (1,1) (1,2) (2,1) (2,2)
        const double n2 = (double)size * (double)size;
        const int la[4] = {1, 1, 2, 2}, lb[4] = {1, 2, 1, 2};
        double v1 = 0.0, v2 = 0.0;
        for (int q = 0; q < 4; q++) {
            const double lg = 5.0 / 3.0 * log10(sqrt((double)(la[q] * la[q] + lb[q] * lb[q])));
            v1 += log10(s[q] / n2) - lg;
            v2 += log10(s[4 + q] / n2) - lg;
        }
        const double c1 = pow(10.0, v1 / 4.0), c2 = pow(10.0, v2 / 4.0);
        if (C1) *C1 = c1;
        if (C2) *C2 = c2;
        e = launch_split_scale(buf, d_out1, d_out2, size * size, 1.0 / sqrt(c1), 1.0 / sqrt(c2), c->stream);
        if (e != hipSuccess) r = fail(c, ATSIM_EHIP, "split_scale: %s", hipGetErrorString(e));
    }
    hipStreamSynchronize(c->stream);
    hipFree(part);
    hipFree(buf);
    return r;
}

int atsim_strfn(atsim_ctx* c, const float* d_screen, long size, long a, long b, double* out) {
    if (!c || !d_screen || !out || size <= 0) return ATSIM_EINVAL;
    const int nblocks = 512;
    double* part = nullptr;
    HIPCHK(c, hipMalloc(&part, sizeof(double) * nblocks));
    long aa = ((a % size) + size) % size;
    hipError_t e = launch_strfn(d_screen, size, aa, b, 1, part, nblocks, c->stream);
    std::vector<double> s;
    int r = e == hipSuccess ? reduce_partials(c, part, nblocks, 1, s) : fail(c, ATSIM_EHIP, "strfn launch");
    hipFree(part);
    if (r) return r;
    *out = s[0] / ((double)size * (double)size);
    return ATSIM_OK;
}

int atsim_r0_measure(atsim_ctx* c, const float* d_screen, long size, long dpix, double* r0) {
    if (!c || !d_screen || !r0 || dpix <= 0 || dpix >= size) return ATSIM_EINVAL;
    const int nblocks = 512;
    double* part = nullptr;
    HIPCHK(c, hipMalloc(&part, sizeof(double) * nblocks));
    hipError_t e = launch_strfn(d_screen, size, dpix, 0, 0, part, nblocks, c->stream);
    std::vector<double> s;
    int r = e == hipSuccess ? reduce_partials(c, part, nblocks, 1, s) : fail(c, ATSIM_EHIP, "r0 launch");
    hipFree(part);
    if (r) return r;
    const double cnt = (double)(size - dpix) * (double)size;
    *r0 = 1.0 * dpix * pow((s[0] / cnt) / 6.88, -3.0 / 5.0);  // AT.c:2269
    return ATSIM_OK;
}

int atsim_scale(atsim_ctx* c, float* d_screen, long n, double scale) {
    if (!c || !d_screen || n <= 0) return ATSIM_EINVAL;
    c->work_gen++;
    HIPCHK(c, launch_scale(d_screen, n, scale, c->stream));
    return ATSIM_OK;
}

// ------------------------------------------------------------------ layers
long atsim_superlayers(const double* alt, const double* cn2, long L, double bin, double* sa, long* sli) {
    if (!alt || !cn2 || !sa || !sli || L <= 0) return ATSIM_EINVAL;
    for (long i = 1; i < L; i++)
        if (alt[i] < alt[i - 1]) return ATSIM_EINVAL;
    std::vector<double> sc(cn2, cn2 + L);
    long S = L;
    for (long i = 0; i < L; i++) sa[i] = alt[i];
    // AT.c:2084-2120: merge the closest pair until every gap exceeds bin
    for (;;) {
        double minaltd = alt[L - 1];
        long index = 0;
        for (long i = 0; i < S - 1; i++) {
            const double v = sa[i + 1] - sa[i];
            if (v < minaltd) { minaltd = v; index = i; }
        }
        if (minaltd > bin || S == 1) break;
        sa[index] = (sc[index] * sa[index] + sc[index + 1] * sa[index + 1]) / (sc[index] + sc[index + 1]);
        sc[index] = sc[index] + sc[index + 1];
        for (long i = index + 1; i < S - 1; i++) { sa[i] = sa[i + 1]; sc[i] = sc[i + 1]; }
        S -= 1;
    }
    // AT.c:2126-2150: separators at mid-points, layer -> super-layer
    const long nsep = S - 1;
    std::vector<double> sep(nsep > 0 ? nsep : 1);
    for (long i = 0; i < nsep; i++) sep[i] = 0.5 * (sa[i] + sa[i + 1]);
    for (long l = 0; l < L; l++) {
        long index = 0;
        for (long i = 0; i + 1 < nsep; i++)
            if (sep[i] < alt[l] && sep[i + 1] > alt[l]) index = i + 1;
        if (nsep > 0 && alt[l] > sep[nsep - 1]) index = nsep;
        sli[l] = index;
    }
    return S;
}

}  // extern "C"

// ------------------------------------------------------------------ simulation
struct Group {
    int lfirst, lcount;  // layers [lfirst, lfirst + lcount)
    double dz;           // propagation distance before this group (0 for the top group)
    int slot0 = 0;       // first phase plane of its layers (groups applied by column passes)
};

struct atsim_sim {
    atsim_ctx* ctx = nullptr;
    long N = 0, Nout = 0, M = 0, L = 0, W = 0, D = 0;
    double dx = 0, dt = 0, zangle = 0;
    int bicubic = 1, fresnel = 1, amplitude = 1;
    std::vector<double> lambda, scoeff, xpos, ypos, vx, vy;
    std::vector<int> unwrap;
    std::vector<Group> groups;  // top -> bottom
    bool propagate = false;     // Fresnel passes run this frame sequence
    float* masters = nullptr;
    // the current frame's buffer set (use_buffers): the fields, the layer
    // positions, the geometric phase's planes and the block sums
    float2* fields = nullptr;
    float2* tf = nullptr;       // [gap][W][N]
    float2* tw = nullptr;
    LayerSample* ls_dev = nullptr;
    LayerSample* ls_host = nullptr;  // pinned ring
    hipEvent_t* ls_ev = nullptr;
    int ls_slot = 0;
    static constexpr int RING = 8;
    double* s_dev = nullptr;
    int* unwrap_dev = nullptr;
    // phase multipliers that factor as s_wl = geo_a[w] geo_b[l]: the planes
    // of the geometric phase's layer sums, made beside the Fresnel passes
    // ([2][D][N][N] FP64: the layers before the last group, then the group)
    bool factored = false;
    double* geo_a = nullptr;  // [W]
    double* geo_b = nullptr;  // [L]
    double* vgeo = nullptr;
    // extra streams of the multi-stream Fresnel passes (set k > 0 on xs_stream[k - 1])
    static constexpr int MAX_SETS = 4;
    hipStream_t xs_stream[MAX_SETS - 1] = {};
    hipEvent_t ev_fork = nullptr, ev_phase = nullptr, ev_first[MAX_SETS - 1] = {}, ev_join[MAX_SETS - 1] = {};
    float* dres = nullptr;   // [F][Nout][Nout] block-rule remainders
    int* blk = nullptr;      // [Nout] block index of each output column / row
    double* bsum = nullptr;   // [F][nb][2] block sums, then rmax in the same allocation
    unsigned* rmax = nullptr; // [F][nb] bits of each block's largest |r| (float)
    int* bstart = nullptr;    // [xs + 1] first output row / column of each block, then Nout
    double* bin = nullptr;
    double* ch = nullptr;
    unsigned* cmax = nullptr;  // [F] largest first-round count of each block-rule field (in bsum's allocation)
    int* flist_dev = nullptr;
    // sampled phase planes of the layers whose screens column passes apply
    // (one pass per gap): [nslots][D][N][N] FP64 quarter turns
    double* vq = nullptr;
    int* slot_layer_dev = nullptr;
    int nslots = 0;
    int nblockfields = 0;
    int xs = 0;
    // per-frame output statistics (atsim_sim_stats_enable): partial sums and
    // the [F][4] sums of pha, pha^2, amp^2, amp^4 of the last frame
    bool stats = false;
    double* st_part = nullptr;
    double* st_out = nullptr;
    // own stream and join event of this simulation inside a multi-simulation frame
    hipStream_t own = nullptr;
    hipEvent_t own_done = nullptr;
    bool masters_borrowed = false;  // masters belong to another simulation (atsim_sim_use_masters)
    // per-launch event timing (bench instrumentation)
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<std::pair<int, size_t>> prof_recs;  // (kind, index of the first event of the pair)
    // captured frame graphs, keyed by the output pointers they were captured
    // with (two sets: double-buffered outputs alternate between them)
    long frames_done = 0;
    unsigned graph_gen = 0;  // the context's tuning generation the captured graphs belong to
    struct GraphSlot {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        float* pha = nullptr;
        float* amp = nullptr;
        void* field = nullptr;
        long used = -1;
        // pipelined frames: exec holds the Fresnel passes, exec2 the
        // extraction and block fix-up; the buffer set they were captured on
        hipGraph_t graph2 = nullptr;
        hipGraphExec_t exec2 = nullptr;
        int parity = 0;
        bool split = false;
    } gslot[4];
    // pipelined frames (tuning key "pipeline"): buffer sets 0 and 1 alternate
    // between frames; set 1 is allocated at the first pipelined frame
    float2* fields_buf[2] = {};
    LayerSample* ls_buf[2] = {};
    double* vgeo_buf[2] = {};
    double* bsum_buf[2] = {};
    int parity = 0;
    hipStream_t hs = nullptr;  // the passes of pipelined frames
    hipEvent_t ev_head[2] = {}, ev_tail[2] = {}, ev_sync = nullptr;
    bool tail_rec[2] = {false, false};
    // the next pipelined frame's passes wait for the context stream: a frame
    // ran there without pipelining, or the simulation's inputs changed
    bool head_sync = true;
    unsigned long seen_gen = 0;
};

enum {
    PROF_ROW = 0,       // row pass: screen + x factor (the first pass), or x factor alone (the last)
    PROF_COL = 1,       // column pass: y factor
    PROF_EXTRACT = 2,
    PROF_BLOCK = 3,
    PROF_ROW_GAP2 = 4,  // row pass: x factor, screen, x factor
    PROF_COL_GAP2 = 5,  // column pass: y factor, screen, y factor
    PROF_PHASE = 6,     // sampled phase planes of the column passes' layers
    PROF_STATS = 7,     // per-frame output statistics
    PROF_GEOMETRY = 8,  // the factored geometric phase's planes
    PROF_KINDS = 9
};

static hipEvent_t prof_event(atsim_sim* s) {
    if (s->ev_used == s->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        s->ev_pool.push_back(e);
    }
    return s->ev_pool[s->ev_used++];
}

// record an event before a launch of `kind`; returns the pair index or -1
static long prof_begin(atsim_sim* s, int kind) {
    if (!s->prof) return -1;
    const size_t i0 = s->ev_used;
    hipEvent_t a = prof_event(s);
    hipEvent_t b = prof_event(s);
    if (!a || !b) return -1;
    (void)hipEventRecord(a, s->ctx->stream);
    s->prof_recs.push_back({kind, i0});
    return (long)i0;
}

static void prof_end(atsim_sim* s, long i0) {
    if (i0 < 0) return;
    (void)hipEventRecord(s->ev_pool[(size_t)i0 + 1], s->ctx->stream);
}

// Destroy a slot's graphs once no launch of them can still be running (the
// frame streams are drained first; a slot is dropped only on a tuning or
// output change, not per frame).
static void drop_graph_slot(atsim_sim* s, atsim_sim::GraphSlot& g) {
    if (!g.exec && !g.exec2) return;
    if (s->ctx) hipStreamSynchronize(s->ctx->stream);
    if (s->hs) hipStreamSynchronize(s->hs);
    if (g.exec) hipGraphExecDestroy(g.exec);
    if (g.graph) hipGraphDestroy(g.graph);
    if (g.exec2) hipGraphExecDestroy(g.exec2);
    if (g.graph2) hipGraphDestroy(g.graph2);
    g = atsim_sim::GraphSlot{};
}

static void sim_drop_graph(atsim_sim* s) {
    for (auto& g : s->gslot) drop_graph_slot(s, g);
}

static void sim_free(atsim_sim* s) {
    if (!s) return;
    if (s->ctx) hipStreamSynchronize(s->ctx->stream);
    if (s->hs) hipStreamSynchronize(s->hs);
    sim_drop_graph(s);
    if (s->own) hipStreamDestroy(s->own);
    if (s->own_done) hipEventDestroy(s->own_done);
    if (!s->masters_borrowed) hipFree(s->masters);
    for (int p = 0; p < 2; p++) {
        hipFree(s->fields_buf[p]);
        hipFree(s->ls_buf[p]);
        hipFree(s->vgeo_buf[p]);
        hipFree(s->bsum_buf[p]);
        if (s->ev_head[p]) hipEventDestroy(s->ev_head[p]);
        if (s->ev_tail[p]) hipEventDestroy(s->ev_tail[p]);
    }
    if (s->ev_sync) hipEventDestroy(s->ev_sync);
    if (s->hs) hipStreamDestroy(s->hs);
    hipFree(s->tf);
    if (s->ls_host) hipHostFree(s->ls_host);
    if (s->ls_ev) {
        for (int i = 0; i < atsim_sim::RING; i++) hipEventDestroy(s->ls_ev[i]);
        delete[] s->ls_ev;
    }
    hipFree(s->s_dev);
    hipFree(s->geo_a);
    hipFree(s->geo_b);
    hipFree(s->unwrap_dev);
    hipFree(s->dres);
    for (int k = 0; k < atsim_sim::MAX_SETS - 1; k++) {
        if (s->xs_stream[k]) hipStreamDestroy(s->xs_stream[k]);
        if (s->ev_first[k]) hipEventDestroy(s->ev_first[k]);
        if (s->ev_join[k]) hipEventDestroy(s->ev_join[k]);
    }
    if (s->ev_fork) hipEventDestroy(s->ev_fork);
    if (s->ev_phase) hipEventDestroy(s->ev_phase);
    hipFree(s->blk);
    hipFree(s->bstart);
    hipFree(s->bin);
    hipFree(s->ch);
    hipFree(s->flist_dev);
    hipFree(s->vq);
    hipFree(s->slot_layer_dev);
    hipFree(s->st_part);
    hipFree(s->st_out);
    for (auto e : s->ev_pool) (void)hipEventDestroy(e);
    delete s;
}

extern "C" {

// doubles of the block-sum allocation, cleared every frame: F nb (cos, sin)
// sums, then F nb unsigned largest |r| and F unsigned largest |count|
static long bsum_doubles(long F, long nb) { return F * nb * 2 + (F * nb + F + 1) / 2; }

// Make buffer set p the current frame's: fields, layer positions, geometric
// phase planes and block sums (with each block's largest |r| and each
// field's largest count in the same allocation).
static void use_buffers(atsim_sim* s, int p) {
    s->parity = p;
    s->fields = s->fields_buf[p];
    s->ls_dev = s->ls_buf[p];
    s->vgeo = s->vgeo_buf[p];
    s->bsum = s->bsum_buf[p];
    s->rmax = nullptr;
    s->cmax = nullptr;
    if (s->bsum) {
        const long F = s->D * s->W, nb = s->xs > 0 ? (long)s->xs * s->xs : 1;
        s->rmax = reinterpret_cast<unsigned*>(s->bsum + F * nb * 2);
        s->cmax = s->rmax + F * nb;
    }
}

// Buffer set 1, the passes' stream and the events of pipelined frames,
// made at the first such frame.
static int pipeline_init(atsim_sim* s) {
    if (s->hs) return ATSIM_OK;
    atsim_ctx* c = s->ctx;
    const long F = s->D * s->W;
    const size_t nn = (size_t)s->N * s->N;
    auto alloc = [&](void** p, size_t bytes) {
        if (*p) return ATSIM_OK;
        if (hipMalloc(p, bytes) != hipSuccess) {
            *p = nullptr;
            return fail(c, ATSIM_ENOMEM, "pipeline: cannot allocate %zu bytes for a second buffer set", bytes);
        }
        return ATSIM_OK;
    };
    if (s->fields_buf[0])
        if (int r = alloc((void**)&s->fields_buf[1], sizeof(float2) * F * nn)) return r;
    if (int r = alloc((void**)&s->ls_buf[1], sizeof(LayerSample) * s->D * s->L)) return r;
    if (s->vgeo_buf[0])
        if (int r = alloc((void**)&s->vgeo_buf[1], sizeof(double) * 2 * s->D * nn)) return r;
    if (s->bsum_buf[0]) {
        const long nb = s->xs > 0 ? (long)s->xs * s->xs : 1;
        if (int r = alloc((void**)&s->bsum_buf[1], sizeof(double) * bsum_doubles(F, nb))) return r;
    }
    for (int p = 0; p < 2; p++) {
        if (!s->ev_head[p]) HIPCHK(c, hipEventCreateWithFlags(&s->ev_head[p], hipEventDisableTiming));
        if (!s->ev_tail[p]) HIPCHK(c, hipEventCreateWithFlags(&s->ev_tail[p], hipEventDisableTiming));
    }
    if (!s->ev_sync) HIPCHK(c, hipEventCreateWithFlags(&s->ev_sync, hipEventDisableTiming));
    HIPCHK(c, hipStreamCreateWithFlags(&s->hs, hipStreamNonBlocking));
    s->head_sync = true;
    return ATSIM_OK;
}

int atsim_sim_create(atsim_ctx* c, const atsim_sim_desc* d, atsim_sim** out) {
    if (!c || !d || !out) return ATSIM_EINVAL;
    *out = nullptr;
    if (!supported_n(d->N)) return fail(c, ATSIM_EUNSUPPORTED, "sim: N=%ld must be a power of two in [16, 8192]", d->N);
    if (d->Nout <= 0 || d->Nout > d->N) return fail(c, ATSIM_EINVAL, "sim: need 0 < Nout <= N");
    if (d->M < d->N) return fail(c, ATSIM_EINVAL, "sim: need N <= M (master %ld, grid %ld)", d->M, d->N);
    if (d->M > (1L << 20)) return fail(c, ATSIM_EINVAL, "sim: master too large");
    if (d->L <= 0 || d->W <= 0 || d->D <= 0) return fail(c, ATSIM_EINVAL, "sim: L, W, D must be > 0");
    if (d->L > 64) return fail(c, ATSIM_EUNSUPPORTED, "sim: at most 64 layers (got %ld)", d->L);
    if (d->D * d->W > 65535) return fail(c, ATSIM_EINVAL, "sim: too many fields");
    if (!(d->dx > 0)) return fail(c, ATSIM_EINVAL, "sim: dx must be > 0");
    if (!d->slayer_alt || !d->super_layer_index || !d->lambda || !d->scoeff || !d->unwrap_mode || !d->xpos0 ||
        !d->ypos0 || !d->vx || !d->vy)
        return fail(c, ATSIM_EINVAL, "sim: null descriptor array");
    for (long l = 1; l < d->L; l++)
        if (d->super_layer_index[l] < d->super_layer_index[l - 1])
            return fail(c, ATSIM_EINVAL, "sim: layers must be in ascending altitude order");
    HIPCHK(c, hipSetDevice(c->device));
    atsim_sim* s = new atsim_sim();
    s->ctx = c;
    s->N = d->N; s->Nout = d->Nout; s->M = d->M; s->L = d->L; s->W = d->W; s->D = d->D;
    s->dx = d->dx; s->dt = d->dt; s->zangle = d->zangle;
    s->bicubic = d->bicubic; s->fresnel = d->fresnel; s->amplitude = d->amplitude;
    s->lambda.assign(d->lambda, d->lambda + d->W);
    s->scoeff.assign(d->scoeff, d->scoeff + d->W * d->L);
    s->unwrap.assign(d->unwrap_mode, d->unwrap_mode + d->W);
    s->xpos.assign(d->xpos0, d->xpos0 + d->D * d->L);
    s->ypos.assign(d->ypos0, d->ypos0 + d->D * d->L);
    s->vx.assign(d->vx, d->vx + d->L);
    s->vy.assign(d->vy, d->vy + d->L);
    // groups in the order of the layer loop, top -> bottom (AT.c:2649-2692)
    for (long l = d->L - 1; l >= 0; l--) {
        if (l == d->L - 1 || d->super_layer_index[l + 1] != d->super_layer_index[l]) {
            Group g;
            g.lfirst = (int)l;
            g.lcount = 1;
            g.dz = (l == d->L - 1) ? 0.0
                                   : (d->slayer_alt[d->super_layer_index[l + 1]] - d->slayer_alt[d->super_layer_index[l]]) /
                                         cos(d->zangle);
            s->groups.push_back(g);
        } else {
            s->groups.back().lfirst = (int)l;
            s->groups.back().lcount++;
        }
    }
    const long G = (long)s->groups.size() - 1;
    s->propagate = s->amplitude && s->fresnel && G > 0;
    const long F = s->D * s->W;
    const size_t nn = (size_t)s->N * s->N, no = (size_t)s->Nout * s->Nout;
    int r = ATSIM_OK;
#define SIMALLOC(ptr, bytes)                                                                       \
    do {                                                                                           \
        if (r == ATSIM_OK && hipMalloc((void**)&(ptr), (bytes)) != hipSuccess)                     \
            r = fail(c, ATSIM_ENOMEM, "sim: cannot allocate %zu bytes for " #ptr, (size_t)(bytes)); \
    } while (0)
    SIMALLOC(s->masters, sizeof(float) * s->L * s->M * s->M);
    if (s->propagate) {
        SIMALLOC(s->fields_buf[0], sizeof(float2) * F * nn);
        SIMALLOC(s->tf, sizeof(float2) * G * s->W * s->N);
    }
    SIMALLOC(s->ls_buf[0], sizeof(LayerSample) * s->D * s->L);
    SIMALLOC(s->s_dev, sizeof(double) * s->W * s->L);
    SIMALLOC(s->unwrap_dev, sizeof(int) * s->W);
    SIMALLOC(s->blk, sizeof(int) * (s->Nout > 0 ? s->Nout : 1));
    // multipliers that factor exactly as s_wl = a_w b_l (a_w = s_w0,
    // b_l = s_0l / s_00; the reference's lambda_ref / lambda rows do, with
    // b_l = 1): the extraction's geometric phase then needs two planes
    std::vector<double> geo_a(s->W), geo_b(s->L);
    s->factored = s->W > 0 && s->L > 0 && s->scoeff[0] != 0.0;
    for (long l = 0; s->factored && l < s->L; l++) geo_b[l] = s->scoeff[l] / s->scoeff[0];
    for (long w = 0; s->factored && w < s->W; w++) {
        geo_a[w] = s->scoeff[w * s->L];
        for (long l = 0; l < s->L; l++)
            if (geo_a[w] * geo_b[l] != s->scoeff[w * s->L + l]) s->factored = false;
    }
    if (s->factored) {
        SIMALLOC(s->geo_a, sizeof(double) * s->W);
        SIMALLOC(s->geo_b, sizeof(double) * s->L);
        SIMALLOC(s->vgeo_buf[0], sizeof(double) * 2 * s->D * nn);
    }
    if (r == ATSIM_OK && hipHostMalloc((void**)&s->ls_host, sizeof(LayerSample) * s->D * s->L * atsim_sim::RING) != hipSuccess)
        r = fail(c, ATSIM_ENOMEM, "sim: pinned allocation failed");
    if (r == ATSIM_OK && (hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming) != hipSuccess ||
                          hipEventCreateWithFlags(&s->ev_phase, hipEventDisableTiming) != hipSuccess))
        r = fail(c, ATSIM_EHIP, "sim: event creation failed");
    for (int k = 0; r == ATSIM_OK && k < atsim_sim::MAX_SETS - 1; k++)
        if (hipStreamCreateWithFlags(&s->xs_stream[k], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&s->ev_first[k], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&s->ev_join[k], hipEventDisableTiming) != hipSuccess)
            r = fail(c, ATSIM_EHIP, "sim: stream / event creation failed");
    if (r == ATSIM_OK) {
        s->ls_ev = new hipEvent_t[atsim_sim::RING];
        for (int i = 0; i < atsim_sim::RING; i++) hipEventCreateWithFlags(&s->ls_ev[i], hipEventDisableTiming);
    }
    std::vector<int> flist;
    for (long dd = 0; dd < s->D; dd++)
        for (long w = 0; w < s->W; w++)
            if (s->unwrap[w] != 0 && s->amplitude) flist.push_back((int)(dd * s->W + w));
    s->nblockfields = (int)flist.size();
    // one pass per gap: column pass k (odd, 1 <= k < G) applies group k
    std::vector<int> slot_layer;
    if (s->propagate && col_screen_supported(s->N) && G >= 2) {
        for (long k = 1; k < G; k += 2) {
            s->groups[k].slot0 = (int)slot_layer.size();
            for (int q = 0; q < s->groups[k].lcount; q++) slot_layer.push_back(s->groups[k].lfirst + q);
        }
    }
    s->nslots = (int)slot_layer.size();
    if (s->nslots > 0) {
        SIMALLOC(s->vq, sizeof(double) * s->nslots * s->D * nn);
        SIMALLOC(s->slot_layer_dev, sizeof(int) * s->nslots);
    }
    s->xs = (int)(s->Nout / 20);
    if (s->nblockfields > 0) {
        const size_t nb = (size_t)(s->xs > 0 ? s->xs * s->xs : 1);
        SIMALLOC(s->dres, sizeof(float) * F * no);
        // block sums and largest |r|, cleared together every frame
        SIMALLOC(s->bsum_buf[0], sizeof(double) * bsum_doubles(F, nb));
        SIMALLOC(s->bstart, sizeof(int) * (s->xs + 1));
        SIMALLOC(s->bin, sizeof(double) * F * nb);
        SIMALLOC(s->ch, sizeof(double) * F * nb);
        SIMALLOC(s->flist_dev, sizeof(int) * flist.size());
    }
#undef SIMALLOC
    use_buffers(s, 0);
    if (r == ATSIM_OK && hipMemsetAsync(s->masters, 0, sizeof(float) * s->L * s->M * s->M, c->stream) != hipSuccess)
        r = fail(c, ATSIM_EHIP, "sim: memset failed");
    if (r == ATSIM_OK) r = get_twiddles(c, s->N, &s->tw);
    if (r == ATSIM_OK && s->propagate) {
        std::vector<float2> h((size_t)G * s->W * s->N);
        for (long g = 1; g <= G; g++)
            for (long w = 0; w < s->W; w++)
                fresnel_table(h.data() + ((g - 1) * s->W + w) * s->N, s->N, s->dx, s->groups[g].dz, s->lambda[w]);
        if (copy_on_stream(c, s->tf, h.data(), sizeof(float2) * h.size()) != hipSuccess)
            r = fail(c, ATSIM_EHIP, "sim: table upload failed");
    }
    // block of output column / row i: (long)(1.0 * i / Nout * xs) (AT.c:3611)
    std::vector<int> blk(s->Nout > 0 ? s->Nout : 1, 0);
    for (long i = 0; i < s->Nout; i++) blk[i] = (int)((double)i / s->Nout * s->xs);
    // blk is non-decreasing in i: block b covers bstart[b] .. bstart[b + 1] - 1
    std::vector<int> bstart((size_t)s->xs + 1, (int)s->Nout);
    for (long i = s->Nout - 1; i >= 0; i--)
        if (blk[i] < s->xs) bstart[blk[i]] = (int)i;
    for (long b = s->xs - 1; b >= 0; b--) bstart[b] = std::min(bstart[b], bstart[b + 1]);
    if (r == ATSIM_OK) {
        if (copy_on_stream(c, s->s_dev, s->scoeff.data(), sizeof(double) * s->scoeff.size()) != hipSuccess ||
            (s->factored && (copy_on_stream(c, s->geo_a, geo_a.data(), sizeof(double) * s->W) != hipSuccess ||
                             copy_on_stream(c, s->geo_b, geo_b.data(), sizeof(double) * s->L) != hipSuccess)) ||
            copy_on_stream(c, s->unwrap_dev, s->unwrap.data(), sizeof(int) * s->W) != hipSuccess ||
            copy_on_stream(c, s->blk, blk.data(), sizeof(int) * blk.size()) != hipSuccess ||
            (s->nblockfields > 0 &&
             (copy_on_stream(c, s->flist_dev, flist.data(), sizeof(int) * flist.size()) != hipSuccess ||
              copy_on_stream(c, s->bstart, bstart.data(), sizeof(int) * bstart.size()) != hipSuccess)) ||
            (s->nslots > 0 &&
             copy_on_stream(c, s->slot_layer_dev, slot_layer.data(), sizeof(int) * slot_layer.size()) != hipSuccess))
            r = fail(c, ATSIM_EHIP, "sim: parameter upload failed");
    }
    if (r == ATSIM_OK && hipStreamSynchronize(c->stream) != hipSuccess) r = fail(c, ATSIM_EHIP, "sim: init sync failed");
    if (r != ATSIM_OK) {
        sim_free(s);
        return r;
    }
    *out = s;
    return ATSIM_OK;
}

void atsim_sim_destroy(atsim_sim* s) { sim_free(s); }

float* atsim_sim_master(atsim_sim* s, long layer) {
    if (!s || layer < 0 || layer >= s->L) return nullptr;
    s->head_sync = true;
    return s->masters + (size_t)layer * s->M * s->M;
}

int atsim_sim_set_master(atsim_sim* s, long layer, const float* src) {
    if (!s || !src || layer < 0 || layer >= s->L) return ATSIM_EINVAL;
    s->head_sync = true;
    s->ctx->work_gen++;
    atsim_ctx* c = s->ctx;
    HIPCHK(c, copy_on_stream(c, s->masters + (size_t)layer * s->M * s->M, src, sizeof(float) * s->M * s->M));
    return ATSIM_OK;
}

int atsim_sim_set_positions(atsim_sim* s, const double* xp, const double* yp) {
    if (!s || !xp || !yp) return ATSIM_EINVAL;
    memcpy(s->xpos.data(), xp, sizeof(double) * s->xpos.size());
    memcpy(s->ypos.data(), yp, sizeof(double) * s->ypos.size());
    return ATSIM_OK;
}

int atsim_sim_positions(atsim_sim* s, double* xp, double* yp) {
    if (!s) return ATSIM_EINVAL;
    if (xp) memcpy(xp, s->xpos.data(), sizeof(double) * s->xpos.size());
    if (yp) memcpy(yp, s->ypos.data(), sizeof(double) * s->ypos.size());
    return ATSIM_OK;
}

long atsim_sim_gaps(atsim_sim* s) { return s ? (long)s->groups.size() - 1 : -1; }

// Catmull-Rom weights for taps -1, 0, 1, 2 (the tensor product is the a_ij
// polynomial of AT.c:2884-2910); bilinear puts 1-f, f on taps 0, 1.
static void tap_weights(double f, int bicubic, double w[4]) {
    if (bicubic) {
        const double f2 = f * f, f3 = f2 * f;
        w[0] = 0.5 * (-f3 + 2.0 * f2 - f);
        w[1] = 0.5 * (3.0 * f3 - 5.0 * f2 + 2.0);
        w[2] = 0.5 * (-3.0 * f3 + 4.0 * f2 + f);
        w[3] = 0.5 * (f3 - f2);
    } else {
        w[0] = 0.0;
        w[1] = 1.0 - f;
        w[2] = f;
        w[3] = 0.0;
    }
}

// one time step of the layer positions (AT.c:2706-2755: cumulative FP64,
// wrapped into [0, M]) and the upload of their integer bases and tap weights
static int advance_positions(atsim_sim* s, hipStream_t st) {
    atsim_ctx* c = s->ctx;
    const long L = s->L, D = s->D, M = s->M;
    LayerSample* slot = s->ls_host + (size_t)s->ls_slot * D * L;
    HIPCHK(c, hipEventSynchronize(s->ls_ev[s->ls_slot]));
    for (long d = 0; d < D; d++)
        for (long l = 0; l < L; l++) {
            double& x = s->xpos[d * L + l];
            double& y = s->ypos[d * L + l];
            x += s->vx[l] * s->dt;
            y += s->vy[l] * s->dt;
            while (x < 0) x += 1.0 * M;
            while (x > 1.0 * M) x -= 1.0 * M;
            while (y < 0) y += 1.0 * M;
            while (y > 1.0 * M) y -= 1.0 * M;
            const double xi = floor(x), yi = floor(y);
            LayerSample& p = slot[d * L + l];
            p.x0 = (int)((long)xi % M);
            p.y0 = (int)((long)yi % M);
            p.pad0 = p.pad1 = 0;
            tap_weights(x - xi, s->bicubic, p.wx);
            tap_weights(y - yi, s->bicubic, p.wy);
        }
    HIPCHK(c, hipMemcpyAsync(s->ls_dev, slot, sizeof(LayerSample) * D * L, hipMemcpyHostToDevice, st));
    HIPCHK(c, hipEventRecord(s->ls_ev[s->ls_slot], st));
    s->ls_slot = (s->ls_slot + 1) % atsim_sim::RING;
    return ATSIM_OK;
}

int atsim_sim_advance(atsim_sim* s) {
    if (!s) return ATSIM_EINVAL;
    s->head_sync = true;
    return advance_positions(s, s->ctx->stream);
}

// Fresnel passes of a frame.  The x and y factors of a gap commute.  With
// one pass per gap, gap g is applied x-first for even g and y-first for odd
// g: pass k (0 .. G; rows for even k, columns for odd k) applies the factor
// of gap k - 1 along its axis, the screen of layer group k and the factor
// of gap k along the same axis -- pass 0 starts from the unit field, pass G
// has no group or gap after it.  With two passes per gap, step 2g is the
// row pass (screen of group g, x factor of gap g) and step 2g + 1 the
// column pass (y factor).
static bool use_gap_fusion(const atsim_sim* s) {
    return s->ctx->gap_passes == 1 && col_screen_supported(s->N) && (long)s->groups.size() - 1 >= 2 &&
           s->nslots > 0;
}

static int step_kind(bool fused, long step, long G) {
    if (!fused) return step % 2 ? PROF_COL : PROF_ROW;
    if (step == 0 || step == G) return step % 2 ? PROF_COL : PROF_ROW;
    return step % 2 ? PROF_COL_GAP2 : PROF_ROW_GAP2;
}

// one step of the chain for the wavelength range of sa (wcnt 0: all)
static hipError_t launch_step(atsim_sim* s, bool fused, long step, ScreenArgs sa, hipStream_t st) {
    atsim_ctx* c = s->ctx;
    const long G = (long)s->groups.size() - 1, W = s->W, N = s->N, D = s->D;
    auto tab = [&](long g) { return (const float2*)(s->tf + (size_t)g * W * N); };
    const long nfields = D * (sa.wcnt > 0 ? sa.wcnt : W);
    auto group = [&](long g) {
        sa.lfirst = s->groups[g].lfirst;
        sa.lcount = s->groups[g].lcount;
    };
    if (!fused) {
        const long g = step / 2;
        group(g);
        if (step % 2 == 0) return launch_row_fresnel(N, g == 0, s->fields, (int)D, tab(g), s->tw, sa, st, c->row_mw);
        return launch_col(N, OP_FRESNEL, 0, s->fields, nfields, tab(g), (int)W, s->tw, st, c->col_c, c->col_e,
                          sa.wlo, sa.wcnt);
    }
    const long k = step;
    const bool rows = k % 2 == 0;
    if (k == 0) {
        group(0);
        return launch_row_fresnel(N, true, s->fields, (int)D, tab(0), s->tw, sa, st, c->row_mw);
    }
    if (k == G) {
        sa.lfirst = 0;
        sa.lcount = 0;
        if (rows) return launch_row_fresnel(N, false, s->fields, (int)D, tab(G - 1), s->tw, sa, st, c->row_mw);
        return launch_col(N, OP_FRESNEL, 0, s->fields, nfields, tab(G - 1), (int)W, s->tw, st, c->col_c, c->col_e,
                          sa.wlo, sa.wcnt);
    }
    group(k);
    if (rows)
        return launch_row_fresnel(N, false, s->fields, (int)D, tab(k), s->tw, sa, st, c->row_mw, tab(k - 1));
    return launch_col_screen(N, s->fields, (int)D, tab(k - 1), tab(k), s->tw, sa, s->vq, s->groups[k].slot0, st,
                             c->col2_e);
}

// the kernels of one frame (positions already uploaded), in stream order on
// ms (the context stream unless a multi-simulation frame gives another)
// Parts of a frame: the Fresnel passes with the planes they read, then the
// extraction, block fix-up and statistics.
enum { FRAME_PASSES = 1, FRAME_EXTRACT = 2, FRAME_ALL = 3 };

static int enqueue_frame(atsim_sim* s, float* d_pha, float* d_amp, void* d_field, hipStream_t ms = nullptr,
                         int part = FRAME_ALL) {
    atsim_ctx* c = s->ctx;
    if (!ms) ms = c->stream;
    const long L = s->L, D = s->D, W = s->W, M = s->M;
    ScreenArgs sa{};
    sa.masters = s->masters;
    sa.ls = s->ls_dev;
    sa.s = s->s_dev;
    sa.M = M;
    sa.L = (int)L;
    sa.W = (int)W;
    const long F = D * W;
    const long G = (long)s->groups.size() - 1;
    const int nsets = (int)std::min<long>(c->streams, W);
    const bool fused = use_gap_fusion(s);
    const long nsteps = G <= 0 ? 0 : (fused ? G + 1 : 2 * G);
    ExtractArgs a{};
    a.fields = s->propagate ? s->fields : nullptr;
    a.field_out = (float2*)d_field;
    a.pha = d_pha;
    a.amp = d_amp;
    a.dres = s->dres;
    RemainderSource rsrc{};
    a.blk = s->blk;
    a.bsum = s->bsum;
    a.rmax = (c->block_apply == 1 && s->xs > 0) ? s->rmax : nullptr;
    a.unwrap_mode = s->unwrap_dev;
    a.all = sa;
    a.all.wlo = a.all.wcnt = 0;
    a.all.lfirst = 0;
    a.all.lcount = (int)L;
    if (s->propagate) {
        a.last_first = s->groups[G].lfirst;
        a.last_count = s->groups[G].lcount;
    } else {
        a.last_first = 0;
        a.last_count = (int)L;
    }
    a.N = (int)s->N;
    a.Nout = (int)s->Nout;
    a.W = (int)W;
    a.D = (int)D;
    a.amplitude = s->amplitude;
    a.xs = s->xs;
    // factored multipliers: the geometric phase's planes, made beside the
    // passes on the second set's stream (or before the extraction on one)
    const bool planes = s->factored && c->extract_planes == 1;
    const bool full = d_field != nullptr;
    if (planes) {
        a.geo_a = s->geo_a;
        a.geo_b = s->geo_b;
        a.vr = s->vgeo;
        a.vl = s->vgeo + (size_t)D * s->N * s->N;
    }
    bool geometry_done = false;
    // the block sums are cleared by the planes sweep when there is one
    const bool clear_sums = s->nblockfields > 0 && s->xs > 0;
    double* const zp = (planes && clear_sums) ? s->bsum : nullptr;
    const long zn = clear_sums ? bsum_doubles(F, (long)s->xs * s->xs) : 0;
    // the sparse final update of the block rule makes the remainders again
    // from the final fields and the planes (the extraction then writes no
    // map); without propagation there are no fields (the field is 1)
    if (planes && !full && a.rmax != nullptr && s->propagate) {
        a.dres = nullptr;
        rsrc = RemainderSource{s->fields, a.vr, a.vl, s->geo_a, (int)s->N, (int)W, (int)(s->N - s->Nout) / 2,
                               a.amplitude && a.last_count > 0};
    }
    auto phase = [&](hipStream_t st) {
        return launch_layer_phase(s->vq, s->masters, s->ls_dev, s->slot_layer_dev, s->nslots, s->N, M, L, D, st);
    };
    if (part & FRAME_PASSES) {
        if (s->propagate && nsets >= 2 && !s->prof) {
            // nsets wavelength sets, each its own chain of passes: set 0 on the
            // context stream, set k on xs_stream[k - 1], started after set
            // k - 1's first pass, so that one set's row pass (arithmetic-bound)
            // runs beside another's column pass (memory-bound).  All join the
            // context stream before extraction.  The phase planes of the column
            // passes are sampled on set 1's stream beside set 0's first (row)
            // pass; every other set's first column pass waits for them.  With
            // factored multipliers the same sweep over the layers makes the
            // extraction's geometric-phase planes.
            HIPCHK(c, hipEventRecord(s->ev_fork, ms));
            for (int k = 1; k < nsets; k++) HIPCHK(c, hipStreamWaitEvent(s->xs_stream[k - 1], s->ev_fork, 0));
            if (fused && planes) {
                // the phase planes and the geometric-phase planes in one sweep
                // over the layers (joins the context stream with set 1's chain)
                HIPCHK(c, launch_extract_geometry(a, full, s->xs_stream[0], s->vq, s->slot_layer_dev, s->nslots, zp, zn));
                HIPCHK(c, hipEventRecord(s->ev_phase, s->xs_stream[0]));
                geometry_done = true;
            } else {
                if (fused) {
                    HIPCHK(c, phase(s->xs_stream[0]));
                    HIPCHK(c, hipEventRecord(s->ev_phase, s->xs_stream[0]));
                }
                if (planes) {
                    HIPCHK(c, launch_extract_geometry(a, full, s->xs_stream[0], nullptr, nullptr, 0, zp, zn));
                    geometry_done = true;
                }
            }
            for (long step = 0; step < nsteps; step++) {
                for (int k = 0; k < nsets; k++) {
                    hipStream_t st = k ? s->xs_stream[k - 1] : ms;
                    sa.wlo = (int)(W * k / nsets);
                    sa.wcnt = (int)(W * (k + 1) / nsets) - sa.wlo;
                    if (fused && step == 1 && k != 1) HIPCHK(c, hipStreamWaitEvent(st, s->ev_phase, 0));
                    HIPCHK(c, launch_step(s, fused, step, sa, st));
                    if (step == 0 && k + 1 < nsets) {
                        HIPCHK(c, hipEventRecord(s->ev_first[k], st));
                        HIPCHK(c, hipStreamWaitEvent(s->xs_stream[k], s->ev_first[k], 0));
                    }
                }
            }
            sa.wlo = sa.wcnt = 0;
            for (int k = 1; k < nsets; k++) {
                HIPCHK(c, hipEventRecord(s->ev_join[k - 1], s->xs_stream[k - 1]));
                HIPCHK(c, hipStreamWaitEvent(ms, s->ev_join[k - 1], 0));
            }
        } else if (s->propagate) {
            if (fused && planes) {
                long pg = prof_begin(s, PROF_GEOMETRY);
                HIPCHK(c, launch_extract_geometry(a, full, ms, s->vq, s->slot_layer_dev, s->nslots, zp, zn));
                prof_end(s, pg);
                geometry_done = true;
            } else if (fused) {
                long pp = prof_begin(s, PROF_PHASE);
                HIPCHK(c, phase(ms));
                prof_end(s, pp);
            }
            for (long step = 0; step < nsteps; step++) {
                long pr = prof_begin(s, step_kind(fused, step, G));
                HIPCHK(c, launch_step(s, fused, step, sa, ms));
                prof_end(s, pr);
            }
        }
        if (planes && !geometry_done) {
            long pg = prof_begin(s, PROF_GEOMETRY);
            HIPCHK(c, launch_extract_geometry(a, full, ms, nullptr, nullptr, 0, zp, zn));
            prof_end(s, pg);
        }
        if (clear_sums && !planes) HIPCHK(c, launch_zero_f64(s->bsum, zn, ms));
    }
    if (!(part & FRAME_EXTRACT)) return ATSIM_OK;
    long pe = prof_begin(s, PROF_EXTRACT);
    HIPCHK(c, launch_extract(a, full, ms));
    prof_end(s, pe);
    if (s->nblockfields > 0) {
        long pb = prof_begin(s, PROF_BLOCK);
        HIPCHK(c, launch_block_fix(s->bsum, s->bin, s->ch, s->cmax, s->flist_dev, s->nblockfields, s->xs, ms));
        if (a.rmax)
            HIPCHK(c, launch_block_apply_sparse(d_pha, a.dres, s->bin, s->rmax, s->flist_dev, s->bstart,
                                                s->nblockfields, (int)s->Nout, s->xs, rsrc, ms));
        else
            HIPCHK(c, launch_block_apply(d_pha, s->dres, s->bin, s->flist_dev, s->blk, s->nblockfields,
                                         (int)s->Nout, s->xs, ms));
        prof_end(s, pb);
    }
    if (s->stats) {
        long ps = prof_begin(s, PROF_STATS);
        HIPCHK(c, launch_stats(d_pha, d_amp, F, s->Nout * s->Nout, s->st_part, s->st_out, ms));
        prof_end(s, ps);
    }
    return ATSIM_OK;
}

// Capture one part of a frame enqueued on st into an instantiated graph.
static int capture_frame(atsim_sim* s, hipStream_t st, int part, float* d_pha, float* d_amp, void* d_field,
                         hipGraph_t* graph, hipGraphExec_t* exec) {
    atsim_ctx* c = s->ctx;
    HIPCHK(c, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int r = enqueue_frame(s, d_pha, d_amp, d_field, st, part);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(st, &g);
    if (r) {
        if (g) hipGraphDestroy(g);
        return r;
    }
    if (e != hipSuccess) return fail(c, ATSIM_EHIP, "frame graph capture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        *exec = nullptr;
        hipGraphDestroy(g);
        return fail(c, ATSIM_EHIP, "frame graph instantiate: %s", hipGetErrorString(e));
    }
    *graph = g;
    return ATSIM_OK;
}

// The captured graphs of a frame with these outputs on the current buffer set
// (split: a pipelined frame's two parts), captured now into the least
// recently used slot when there are none.
static int frame_graphs(atsim_sim* s, float* d_pha, float* d_amp, void* d_field, bool split,
                        atsim_sim::GraphSlot** out) {
    atsim_ctx* c = s->ctx;
    atsim_sim::GraphSlot* gs = nullptr;
    for (auto& g : s->gslot)
        if (g.exec && g.pha == d_pha && g.amp == d_amp && g.field == d_field && g.parity == s->parity &&
            g.split == split)
            gs = &g;
    if (!gs) {
        gs = &s->gslot[0];
        for (auto& g : s->gslot)
            if (g.used < gs->used) gs = &g;
        drop_graph_slot(s, *gs);
        gs->used = -1;
        if (split) {
            if (int r = capture_frame(s, s->hs, FRAME_PASSES, d_pha, d_amp, d_field, &gs->graph, &gs->exec)) return r;
            if (int r = capture_frame(s, c->stream, FRAME_EXTRACT, d_pha, d_amp, d_field, &gs->graph2, &gs->exec2)) {
                drop_graph_slot(s, *gs);
                return r;
            }
        } else if (int r = capture_frame(s, c->stream, FRAME_ALL, d_pha, d_amp, d_field, &gs->graph, &gs->exec)) {
            return r;
        }
        gs->pha = d_pha;
        gs->amp = d_amp;
        gs->field = d_field;
        gs->parity = s->parity;
        gs->split = split;
    }
    gs->used = s->frames_done;
    *out = gs;
    return ATSIM_OK;
}

int atsim_sim_frame(atsim_sim* s, float* d_pha, float* d_amp, void* d_field) {
    if (!s || !d_pha || !d_amp) return ATSIM_EINVAL;
    atsim_ctx* c = s->ctx;
    // Graph mode replays captured kernel sequences from the second frame on
    // (the first frame sets the kernels' LDS attributes); per-launch
    // profiling needs the eager path.  After a tuning change the graphs are
    // dropped, this frame runs eagerly (first launches of other kernel
    // variants set their LDS attributes outside a capture) and the next one
    // captures again.
    bool retuned = false;
    if (s->graph_gen != c->tuning_gen) {
        sim_drop_graph(s);
        s->graph_gen = c->tuning_gen;
        retuned = true;
    }
    const bool eager = !c->graph || s->prof || s->frames_done == 0 || retuned;
    const bool pipe = c->pipeline == 1 && s->propagate && !s->prof;
    if (!pipe) {
        // the whole frame on the context stream; a later pipelined frame's
        // passes wait for it
        s->head_sync = true;
        if (int r = advance_positions(s, c->stream)) return r;
        if (eager) {
            s->frames_done++;
            return enqueue_frame(s, d_pha, d_amp, d_field);
        }
        atsim_sim::GraphSlot* gs = nullptr;
        if (int r = frame_graphs(s, d_pha, d_amp, d_field, false, &gs)) return r;
        HIPCHK(c, hipGraphLaunch(gs->exec, c->stream));
        s->frames_done++;
        return ATSIM_OK;
    }
    // Pipelined: the passes on the simulation's stream hs into buffer set p,
    // the extraction and fix-up on the context stream after them.  Set p was
    // last read by the extraction of two frames back, so the passes wait for
    // it; they wait for the whole context stream when something other than
    // this simulation's frames wrote device memory there since its last frame.
    if (int r = pipeline_init(s)) return r;
    const int p = s->parity ^ 1;
    use_buffers(s, p);
    hipStream_t hs = s->hs, ts = c->stream;
    if (s->tail_rec[p]) HIPCHK(c, hipStreamWaitEvent(hs, s->ev_tail[p], 0));
    if (s->head_sync || s->seen_gen != c->work_gen) {
        HIPCHK(c, hipEventRecord(s->ev_sync, ts));
        HIPCHK(c, hipStreamWaitEvent(hs, s->ev_sync, 0));
        s->head_sync = false;
        s->seen_gen = c->work_gen;
    }
    if (int r = advance_positions(s, hs)) return r;
    atsim_sim::GraphSlot* gs = nullptr;
    if (!eager)
        if (int r = frame_graphs(s, d_pha, d_amp, d_field, true, &gs)) return r;
    if (gs) HIPCHK(c, hipGraphLaunch(gs->exec, hs));
    else if (int r = enqueue_frame(s, d_pha, d_amp, d_field, hs, FRAME_PASSES)) return r;
    HIPCHK(c, hipEventRecord(s->ev_head[p], hs));
    HIPCHK(c, hipStreamWaitEvent(ts, s->ev_head[p], 0));
    if (gs) HIPCHK(c, hipGraphLaunch(gs->exec2, ts));
    else if (int r = enqueue_frame(s, d_pha, d_amp, d_field, ts, FRAME_EXTRACT)) return r;
    HIPCHK(c, hipEventRecord(s->ev_tail[p], ts));
    s->tail_rec[p] = true;
    s->frames_done++;
    return ATSIM_OK;
}

int atsim_sims_frame(atsim_sim** sims, long n, float** d_pha, float** d_amp) {
    if (!sims || n <= 0 || !d_pha || !d_amp) return ATSIM_EINVAL;
    atsim_ctx* c = sims[0] ? sims[0]->ctx : nullptr;
    if (!c) return ATSIM_EINVAL;
    for (long k = 0; k < n; k++) {
        if (!sims[k] || sims[k]->ctx != c || !d_pha[k] || !d_amp[k])
            return fail(c, ATSIM_EINVAL, "sims_frame: simulation %ld is null, has no outputs or another context", k);
        for (long q = 0; q < k; q++)
            if (sims[q] == sims[k]) return fail(c, ATSIM_EINVAL, "sims_frame: simulation %ld appears twice", k);
        if (sims[k]->prof) return fail(c, ATSIM_EINVAL, "sims_frame: per-launch profiling is on for simulation %ld", k);
    }
    for (long k = 0; k < n; k++) {
        atsim_sim* s = sims[k];
        if (!s->own) {
            HIPCHK(c, hipStreamCreateWithFlags(&s->own, hipStreamNonBlocking));
            HIPCHK(c, hipEventCreateWithFlags(&s->own_done, hipEventDisableTiming));
        }
    }
    if (!c->mg_fork) HIPCHK(c, hipEventCreateWithFlags(&c->mg_fork, hipEventDisableTiming));
    for (long k = 0; k < n; k++)
        if (int r = advance_positions(sims[k], c->stream)) return r;
    // every simulation's frame on its own stream, forked from and joined to
    // the context stream: frames of different grid sizes run side by side.
    // The launches are eager (this path captures no graph).
    HIPCHK(c, hipEventRecord(c->mg_fork, c->stream));
    for (long k = 0; k < n; k++) {
        atsim_sim* s = sims[k];
        s->head_sync = true;  // a later pipelined frame's passes wait for this one
        HIPCHK(c, hipStreamWaitEvent(s->own, c->mg_fork, 0));
        if (int r = enqueue_frame(s, d_pha[k], d_amp[k], nullptr, s->own)) return r;
        HIPCHK(c, hipEventRecord(s->own_done, s->own));
        HIPCHK(c, hipStreamWaitEvent(c->stream, s->own_done, 0));
    }
    c->mg_frames++;
    return ATSIM_OK;
}

int atsim_sim_use_masters(atsim_sim* s, const atsim_sim* owner) {
    if (!s || !owner || s == owner) return ATSIM_EINVAL;
    s->head_sync = true;
    s->ctx->work_gen++;
    atsim_ctx* c = s->ctx;
    if (owner->ctx != c || owner->M != s->M || owner->L != s->L)
        return fail(c, ATSIM_EINVAL, "use_masters: the owner needs the same context, M and L");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (!s->masters_borrowed) hipFree(s->masters);
    s->masters = owner->masters;
    s->masters_borrowed = true;
    sim_drop_graph(s);
    return ATSIM_OK;
}

int atsim_sim_synth_masters(atsim_sim* s, uint64_t seed, const double* outerscale_px, const double* innerscale_px) {
    if (!s || !outerscale_px || !innerscale_px) return ATSIM_EINVAL;
    s->head_sync = true;
    s->ctx->work_gen++;
    atsim_ctx* c = s->ctx;
    if (!supported_n(s->M)) return fail(c, ATSIM_EUNSUPPORTED, "synth_masters: M=%ld unsupported", s->M);
    float2* noise = nullptr;
    HIPCHK(c, hipMalloc(&noise, sizeof(float2) * s->M * s->M));
    int r = ATSIM_OK;
    for (long l = 0; l < s->L && r == ATSIM_OK; l++) {
        r = atsim_noise_c64(c, noise, s->M * s->M, seed, (uint32_t)l);
        if (r == ATSIM_OK)
            r = atsim_make_master(c, noise, s->M, outerscale_px[l], innerscale_px[l], 0.0,
                                  s->masters + (size_t)l * s->M * s->M, nullptr, nullptr, nullptr);
    }
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(noise);
    return r;
}

int atsim_sim_normalize_masters(atsim_sim* s, double r0_goal_px, const double* cn2, double zangle,
                                double* r0_measured) {
    if (!s || !cn2 || !(r0_goal_px > 0)) return ATSIM_EINVAL;
    s->head_sync = true;
    s->ctx->work_gen++;
    atsim_ctx* c = s->ctx;
    const long dpix = 50 < s->M / 2 ? 50 : s->M / 4;
    double r0tot = 0.0;
    for (long l = 0; l < s->L; l++) {
        double r0 = 0.0;
        if (int r = atsim_r0_measure(c, s->masters + (size_t)l * s->M * s->M, s->M, dpix, &r0)) return r;
        if (r0_measured) r0_measured[l] = r0;
        r0tot += r0;
    }
    // AT.c:2271-2285: one factor for all masters so the mean r0 hits the goal
    const double tmp1 = pow((r0tot / s->L) / r0_goal_px, 5.0 / 6.0);
    for (long l = 0; l < s->L; l++) {
        float* m = s->masters + (size_t)l * s->M * s->M;
        if (int r = atsim_scale(c, m, s->M * s->M, tmp1)) return r;
        // AT.c:2339-2351: layer weight sqrt(cn2 / cos z)
        if (int r = atsim_scale(c, m, s->M * s->M, sqrt(cn2[l] / cos(zangle)))) return r;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ATSIM_OK;
}

// ------------------------------------------------------------------ autotuner
// Launch-shape search for one simulation: coordinate descent over the tuning
// keys (two sweeps), each candidate timed over `frames` frames with events on
// the context stream after two untimed frames (the first runs eagerly after
// the change, the second captures the frame graph when graphs are on).
namespace {
struct TuneCfg {
    int row_mw, col_c, col_e, col2_e, streams, gap_passes;
};
TuneCfg tune_get(const atsim_ctx* c) { return {c->row_mw, c->col_c, c->col_e, c->col2_e, c->streams, c->gap_passes}; }
void tune_set(atsim_ctx* c, const TuneCfg& t) {
    c->row_mw = t.row_mw;
    c->col_c = t.col_c;
    c->col_e = t.col_e;
    c->col2_e = t.col2_e;
    c->streams = t.streams;
    c->gap_passes = t.gap_passes;
    c->tuning_gen++;
}
int& tune_field(TuneCfg& t, int k) {
    switch (k) {
        case 0: return t.gap_passes;
        case 1: return t.streams;
        case 2: return t.row_mw;
        case 3: return t.col2_e;
        case 4: return t.col_c;
        default: return t.col_e;
    }
}
const char* tune_name(int k) {
    static const char* names[] = {"gap_passes", "streams", "row_mw", "col2_e", "col_c", "col_e"};
    return names[k];
}
}  // namespace

int atsim_sim_autotune(atsim_sim* s, float* d_pha, float* d_amp, int frames, char* report, long report_len) {
    if (!s || !d_pha || !d_amp || frames <= 0 || frames > 1000) return ATSIM_EINVAL;
    s->head_sync = true;
    atsim_ctx* c = s->ctx;
    if (s->prof) return fail(c, ATSIM_EINVAL, "autotune: per-launch profiling is on");
    const long N = s->N, W = s->W;
    std::vector<double> xp((size_t)(s->D * s->L)), yp((size_t)(s->D * s->L));
    xp.assign(s->xpos.begin(), s->xpos.end());
    yp.assign(s->ypos.begin(), s->ypos.end());
    // candidate values per key (only shapes the kernels support at this N)
    std::vector<int> cand[6];
    cand[0] = {1, 2};
    for (int v = 1; v <= 4 && v <= W; v++) cand[1].push_back(v);
    const long mwmax = std::min<long>(8, std::max<long>(1, 16384 / N));
    for (int v = 1; v <= mwmax; v *= 2) cand[2].push_back(v);
    cand[3] = (N >= 256 && N <= 2048) ? std::vector<int>{16, 32} : std::vector<int>{16};
    for (int v : {2, 4, 8, 16})
        if ((size_t)v * (N + N / 16) * 8 <= 160 * 1024 && v <= N) cand[4].push_back(v);
    if (cand[4].empty()) cand[4].push_back(0);
    cand[5] = (N >= 256 && N <= 2048) ? std::vector<int>{16, 32} : std::vector<int>{16};
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        hipEventDestroy(e0);
        return fail(c, ATSIM_EHIP, "autotune: event creation failed");
    }
    const TuneCfg start = tune_get(c);
    std::string rep;
    int r = ATSIM_OK;
    auto measure = [&](const TuneCfg& t, double* ms) -> int {
        tune_set(c, t);
        for (int k = 0; k < 2; k++)
            if (int rr = atsim_sim_frame(s, d_pha, d_amp, nullptr)) return rr;
        if (hipEventRecord(e0, c->stream) != hipSuccess) return fail(c, ATSIM_EHIP, "autotune: event record");
        for (int k = 0; k < frames; k++)
            if (int rr = atsim_sim_frame(s, d_pha, d_amp, nullptr)) return rr;
        if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
            return fail(c, ATSIM_EHIP, "autotune: event wait");
        float el = 0.f;
        if (hipEventElapsedTime(&el, e0, e1) != hipSuccess) return fail(c, ATSIM_EHIP, "autotune: elapsed");
        *ms = el / frames;
        char line[160];
        snprintf(line, sizeof line, "gap_passes=%d streams=%d row_mw=%d col2_e=%d col_c=%d col_e=%d ms=%.4f\n",
                 t.gap_passes, t.streams, t.row_mw, t.col2_e, t.col_c, t.col_e, *ms);
        rep += line;
        return ATSIM_OK;
    };
    TuneCfg best = start;
    double best_ms = 0.0;
    r = measure(best, &best_ms);
    for (int sweep = 0; sweep < 2 && r == ATSIM_OK; sweep++) {
        bool changed = false;
        for (int k = 0; k < 6 && r == ATSIM_OK; k++) {
            for (int v : cand[k]) {
                TuneCfg t = best;
                if (tune_field(t, k) == v) continue;
                tune_field(t, k) = v;
                double ms = 0.0;
                r = measure(t, &ms);
                if (r) break;
                // a new shape has to beat the current one by more than the
                // timing noise of a few frames (1 %)
                if (ms < 0.99 * best_ms) {
                    best_ms = ms;
                    best = t;
                    changed = true;
                }
            }
        }
        if (!changed) break;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    tune_set(c, r == ATSIM_OK ? best : start);
    // the simulation continues from where it was before the search
    for (size_t i = 0; i < xp.size(); i++) {
        s->xpos[i] = xp[i];
        s->ypos[i] = yp[i];
    }
    if (r == ATSIM_OK) {
        char line[200];
        snprintf(line, sizeof line, "best gap_passes=%d streams=%d row_mw=%d col2_e=%d col_c=%d col_e=%d ms=%.4f\n",
                 best.gap_passes, best.streams, best.row_mw, best.col2_e, best.col_c, best.col_e, best_ms);
        rep += line;
    }
    if (report && report_len > 0) {
        strncpy(report, rep.c_str(), (size_t)report_len - 1);
        report[report_len - 1] = '\0';
    }
    return r;
}

int atsim_sim_profile(atsim_sim* s, int enable) {
    if (!s) return ATSIM_EINVAL;
    s->head_sync = true;
    HIPCHK(s->ctx, hipStreamSynchronize(s->ctx->stream));
    s->prof = enable != 0;
    s->ev_used = 0;
    s->prof_recs.clear();
    return ATSIM_OK;
}

int atsim_sim_profile_read(atsim_sim* s, int kind, double* ms, long* count) {
    if (!s || kind < 0 || kind >= PROF_KINDS) return ATSIM_EINVAL;
    atsim_ctx* c = s->ctx;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double tot = 0.0;
    long n = 0;
    for (auto& r : s->prof_recs) {
        if (r.first != kind) continue;
        float t = 0.f;
        HIPCHK(c, hipEventElapsedTime(&t, s->ev_pool[r.second], s->ev_pool[r.second + 1]));
        tot += t;
        n++;
    }
    if (ms) *ms = tot;
    if (count) *count = n;
    return ATSIM_OK;
}

// ------------------------------------------------------------------ layer screens
int atsim_sim_block_sums(atsim_sim* s, double* out, long n) {
    if (!s || !out) return ATSIM_EINVAL;
    s->head_sync = true;
    atsim_ctx* c = s->ctx;
    const long F = s->D * s->W, need = F * s->xs * s->xs * 2;
    if (n != need) return fail(c, ATSIM_EINVAL, "block sums: n = %ld, expected %ld", n, need);
    if (need == 0) return ATSIM_OK;
    if (s->nblockfields == 0 || !s->bsum) {
        memset(out, 0, sizeof(double) * need);
        return ATSIM_OK;
    }
    HIPCHK(c, copy_on_stream(c, out, s->bsum, sizeof(double) * need));
    return ATSIM_OK;
}

int atsim_sim_stats_enable(atsim_sim* s, int enable) {
    if (!s) return ATSIM_EINVAL;
    s->head_sync = true;
    atsim_ctx* c = s->ctx;
    const long F = s->D * s->W;
    if (enable && !s->st_out) {
        HIPCHK(c, hipMalloc(&s->st_part, sizeof(double) * F * STATS_PART_PER_FIELD));
        HIPCHK(c, hipMalloc(&s->st_out, sizeof(double) * F * 4));
        HIPCHK(c, hipMemsetAsync(s->st_out, 0, sizeof(double) * F * 4, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    if ((enable != 0) != s->stats) sim_drop_graph(s);  // the captured frames change
    s->stats = enable != 0;
    return ATSIM_OK;
}

int atsim_sim_stats(atsim_sim* s, double* out, long n) {
    if (!s || !out) return ATSIM_EINVAL;
    atsim_ctx* c = s->ctx;
    const long F = s->D * s->W;
    if (n < F * 4) return fail(c, ATSIM_EINVAL, "sim_stats: need %ld values, got %ld", F * 4, n);
    if (!s->st_out) return fail(c, ATSIM_EINVAL, "sim_stats: statistics are not enabled");
    std::vector<double> sums((size_t)F * 4);
    HIPCHK(c, hipMemcpyAsync(sums.data(), s->st_out, sizeof(double) * F * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const double np = (double)s->Nout * (double)s->Nout;
    for (long f = 0; f < F; f++) {
        const double lam = s->lambda[f % s->W];
        const double mp = sums[f * 4] / np;
        const double vp = std::max(sums[f * 4 + 1] / np - mp * mp, 0.0);
        const double mi = sums[f * 4 + 2] / np;
        const double mi2 = sums[f * 4 + 3] / np;
        out[f * 4 + 0] = mp * lam / (2.0 * M_PI);
        out[f * 4 + 1] = sqrt(vp) * lam / (2.0 * M_PI);
        out[f * 4 + 2] = mi;
        out[f * 4 + 3] = mi > 0.0 ? mi2 / (mi * mi) - 1.0 : 0.0;
    }
    return ATSIM_OK;
}

int atsim_block_fixup(atsim_ctx* c, const double* bsum, long xs, double* peaks) {
    if (!c || !bsum || !peaks || xs <= 0 || xs > 4096) return ATSIM_EINVAL;
    c->work_gen++;
    const size_t nb = (size_t)xs * xs;
    double *d_sum = nullptr, *d_bin = nullptr, *d_ch = nullptr;
    unsigned* d_max = nullptr;
    int* d_f = nullptr;
    int r = ATSIM_OK;
    if (hipMalloc(&d_sum, sizeof(double) * nb * 2) != hipSuccess || hipMalloc(&d_bin, sizeof(double) * nb) != hipSuccess ||
        hipMalloc(&d_ch, sizeof(double) * nb) != hipSuccess || hipMalloc(&d_max, sizeof(unsigned)) != hipSuccess ||
        hipMalloc(&d_f, sizeof(int)) != hipSuccess)
        r = fail(c, ATSIM_ENOMEM, "block fix-up: out of device memory");
    const int f0 = 0;
    const unsigned z0 = 0u;
    if (r == ATSIM_OK && (copy_on_stream(c, d_sum, bsum, sizeof(double) * nb * 2) != hipSuccess ||
                          copy_on_stream(c, d_f, &f0, sizeof(int)) != hipSuccess ||
                          copy_on_stream(c, d_max, &z0, sizeof(unsigned)) != hipSuccess))
        r = fail(c, ATSIM_EHIP, "block fix-up: upload failed");
    if (r == ATSIM_OK && launch_block_fix(d_sum, d_bin, d_ch, d_max, d_f, 1, (int)xs, c->stream) != hipSuccess)
        r = fail(c, ATSIM_EHIP, "block fix-up: launch failed");
    if (r == ATSIM_OK && copy_on_stream(c, peaks, d_bin, sizeof(double) * nb) != hipSuccess)
        r = fail(c, ATSIM_EHIP, "block fix-up: download failed");
    (void)hipStreamSynchronize(c->stream);
    hipFree(d_sum);
    hipFree(d_bin);
    hipFree(d_ch);
    hipFree(d_max);
    hipFree(d_f);
    return r;
}

int atsim_block_counts(atsim_ctx* c, const double* peaks, long xs, float* counts) {
    if (!c || !peaks || !counts || xs <= 0 || xs > 4096) return ATSIM_EINVAL;
    c->work_gen++;
    const size_t nb = (size_t)xs * xs;
    double* d_bin = nullptr;
    float* d_ch = nullptr;
    unsigned* d_max = nullptr;
    int* d_f = nullptr;
    int r = ATSIM_OK;
    if (hipMalloc(&d_bin, sizeof(double) * nb) != hipSuccess || hipMalloc(&d_ch, sizeof(float) * nb) != hipSuccess ||
        hipMalloc(&d_max, sizeof(unsigned)) != hipSuccess || hipMalloc(&d_f, sizeof(int)) != hipSuccess)
        r = fail(c, ATSIM_ENOMEM, "block counts: out of device memory");
    const int f0 = 0;
    const unsigned z0 = 0u;
    if (r == ATSIM_OK && (copy_on_stream(c, d_bin, peaks, sizeof(double) * nb) != hipSuccess ||
                          copy_on_stream(c, d_f, &f0, sizeof(int)) != hipSuccess ||
                          copy_on_stream(c, d_max, &z0, sizeof(unsigned)) != hipSuccess))
        r = fail(c, ATSIM_EHIP, "block counts: upload failed");
    if (r == ATSIM_OK && launch_block_counts(d_bin, d_ch, d_max, d_f, 1, (int)xs, c->stream) != hipSuccess)
        r = fail(c, ATSIM_EHIP, "block counts: launch failed");
    if (r == ATSIM_OK && copy_on_stream(c, counts, d_ch, sizeof(float) * nb) != hipSuccess)
        r = fail(c, ATSIM_EHIP, "block counts: download failed");
    (void)hipStreamSynchronize(c->stream);
    hipFree(d_bin);
    hipFree(d_ch);
    hipFree(d_max);
    hipFree(d_f);
    return r;
}

int atsim_sim_layer_screens(atsim_sim* s, float* d_out) {
    if (!s || !d_out) return ATSIM_EINVAL;
    s->head_sync = true;
    atsim_ctx* c = s->ctx;
    HIPCHK(c, launch_layer_screens(s->masters, s->ls_dev, d_out, s->N, s->M, s->L, s->D, c->stream));
    return ATSIM_OK;
}

// ------------------------------------------------------------------ surrogate network
int atsim_gemm_bf16(atsim_ctx* c, long M, long N, long K, long batch, float alpha, const float* A, long sab,
                    long sam, long sak, const float* B, long sbb, long sbk, long sbn, float beta, float* C, long scb,
                    long scm, long scn) {
    return atsim_gemm_bf16_ex(c, M, N, K, batch, alpha, A, sab, sam, sak, B, sbb, sbk, sbn, beta, C, scb, scm, scn,
                              nullptr, nullptr, 0);
}

int atsim_gemm_bf16_ex(atsim_ctx* c, long M, long N, long K, long batch, float alpha, const float* A, long sab,
                       long sam, long sak, const float* B, long sbb, long sbk, long sbn, float beta, float* C,
                       long scb, long scm, long scn, const float* bias, float* pre, int gelu) {
    if (!c || !A || !B || !C || M < 0 || N < 0 || K < 0 || batch < 0) return ATSIM_EINVAL;
    c->work_gen++;
    if (sab < 0 || sam < 0 || sak < 0 || sbb < 0 || sbk < 0 || sbn < 0 || scb < 0 || scm < 0 || scn < 0)
        return fail(c, ATSIM_EINVAL, "gemm: negative stride");
    GemmArgs g{M, N, K, batch, alpha, beta, A, sab, sam, sak, B, sbb, sbk, sbn, C, scb, scm, scn, bias, pre,
               gelu ? 1 : 0, 0, 0};
    hipError_t e = launch_gemm_bf16(g, c->stream);
    if (e == hipErrorInvalidValue) return fail(c, ATSIM_EUNSUPPORTED, "gemm: grid too large (M %ld, batch %ld)", M, batch);
    HIPCHK(c, e);
    return ATSIM_OK;
}

int atsim_nn_add_bias_act(atsim_ctx* c, const float* a, const float* b, const float* bias, float* pre, float* y,
                          long C, long P, long n, int gelu) {
    if (!c || !a || !y || C <= 0 || P <= 0 || n < 0 || n % (C * P)) return ATSIM_EINVAL;
    c->work_gen++;
    HIPCHK(c, launch_add_bias_act(a, b, bias, pre, y, C, P, n, gelu, c->stream));
    return ATSIM_OK;
}

int atsim_nn_gelu_bwd(atsim_ctx* c, const float* pre, const float* dy, float* dx, long n) {
    if (!c || !pre || !dy || !dx || n < 0) return ATSIM_EINVAL;
    c->work_gen++;
    HIPCHK(c, launch_gelu_bwd(pre, dy, dx, n, c->stream));
    return ATSIM_OK;
}

int atsim_nn_bias_grad(atsim_ctx* c, const float* dy, float* db, long S, long C, long P, int accumulate) {
    if (!c || !dy || !db || S <= 0 || C <= 0 || P <= 0) return ATSIM_EINVAL;
    c->work_gen++;
    const int nchunk = 64;
    const size_t need = sizeof(float) * nchunk * C;
    if (c->nn_bias_cap < need) {
        if (c->nn_bias) hipFree(c->nn_bias);
        c->nn_bias = nullptr;
        c->nn_bias_cap = 0;
        HIPCHK(c, hipMalloc(&c->nn_bias, need));
        c->nn_bias_cap = need;
    }
    HIPCHK(c, launch_bias_grad(dy, c->nn_bias, S, C, P, nchunk, c->stream));
    HIPCHK(c, launch_reduce_batches(c->nn_bias, db, nchunk, C, accumulate, c->stream));
    return ATSIM_OK;
}

int atsim_nn_mse(atsim_ctx* c, const float* pred, const float* target, float* grad, long n, double* loss) {
    if (!c || !pred || !target || n <= 0) return ATSIM_EINVAL;
    c->work_gen++;
    const int nblk = 256;
    if (!c->nn_part) HIPCHK(c, hipMalloc(&c->nn_part, sizeof(double) * nblk));
    HIPCHK(c, launch_mse(pred, target, grad, n, c->nn_part, nblk, c->stream));
    if (loss) {
        std::vector<double> v;
        int r = reduce_partials(c, c->nn_part, nblk, 1, v);
        if (r) return r;
        *loss = v[0] / (double)n;
    }
    return ATSIM_OK;
}

int atsim_nn_reduce_batches(atsim_ctx* c, const float* part, float* out, long nb, long n, int accumulate) {
    if (!c || !part || !out || nb <= 0 || n <= 0) return ATSIM_EINVAL;
    c->work_gen++;
    HIPCHK(c, launch_reduce_batches(part, out, nb, n, accumulate, c->stream));
    return ATSIM_OK;
}

int atsim_nn_make_pairs(atsim_ctx* c, const float* screens, const double* d_scoeff, const float* pha,
                        const float* amp, long D, long L, long N, float* x, float* t, long Stot, long s0, double xs,
                        double xl, double cs, double ps) {
    if (!c || !screens || !d_scoeff || !x || D <= 0 || L <= 0 || N <= 1 || s0 < 0 || s0 + D > Stot) return ATSIM_EINVAL;
    c->work_gen++;
    if (!(xs > 0) || !(xl > 0) || !(cs > 0) || !(ps > 0)) return fail(c, ATSIM_EINVAL, "make_pairs: scales must be > 0");
    HIPCHK(c, launch_nn_pairs(screens, d_scoeff, pha, amp, D, L, N, Stot, s0, x, t, xs, xl, cs, ps, c->stream));
    return ATSIM_OK;
}

int atsim_nn_unpack(atsim_ctx* c, const float* y, const float* x, long L, long S, long P, double xs, double cs,
                    double ps, float* amp, float* pha) {
    if (!c || !y || !x || !amp || !pha || L <= 0 || S <= 0 || P <= 0) return ATSIM_EINVAL;
    c->work_gen++;
    HIPCHK(c, launch_nn_unpack(y, x, L, S, P, xs, cs, ps, amp, pha, c->stream));
    return ATSIM_OK;
}

int atsim_nn_adam(atsim_ctx* c, float* w, const float* g, float* m, float* v, long n, float lr, float b1, float b2,
                  float eps, long step) {
    if (!c || !w || !g || !m || !v || n < 0 || step < 1) return ATSIM_EINVAL;
    c->work_gen++;
    const float c1 = (float)(1.0 - pow((double)b1, (double)step)), c2 = (float)(1.0 - pow((double)b2, (double)step));
    HIPCHK(c, launch_adam(w, g, m, v, n, lr, b1, b2, eps, c1, c2, c->stream));
    return ATSIM_OK;
}

}  // extern "C"
