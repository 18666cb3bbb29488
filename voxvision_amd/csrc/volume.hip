// Frames and volumes on either side of the ray-casting path:
//   * voxg_redraw: the square-to-row copy of vox_redraw
//     (src/voxrnd/renderer.c:321-327, src/voxrnd/copy-helper.c:5-40);
//   * voxg_raw_select / voxg_read_raw_data: the sample loop of
//     vox_read_raw_data (src/voxtrees/datareader.c:11-85) as an ordered
//     compaction on the GPU, followed by the host tree builder.
// Both are byte moves bound by HBM bandwidth.
#include <hip/hip_runtime.h>

#include <errno.h>
#include <fcntl.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdint>
#include <string>

#include "../../include/voxgpu.h"
#include "voxg_internal.h"

using voxg_internal::fail;

namespace {

// ---------------------------------------------------------------------------
// copy_squares: thread t moves line (t & 3) of square (t >> 2), 16 bytes
// in and 16 bytes out.  A wave reads 1 KiB contiguously and writes four
// 256-byte row segments.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
redraw_kernel(uint4 *__restrict__ sq, uint4 *__restrict__ px, uint32_t ws,
              uint32_t bpf, uint32_t W, uint64_t nthreads, int clear) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nthreads) return;
    const uint64_t s = t >> 2;                 // square, all frames
    const uint32_t line = (uint32_t)(t & 3u);  // pixel row in the square
    const uint64_t f = s / bpf;
    const uint32_t cs = (uint32_t)(s - f * bpf);
    const uint32_t sx = cs % ws, sy = cs / ws;
    const uint4 v = sq[t];
    const uint64_t pix = f * (uint64_t)bpf * 16u +
                         (uint64_t)(4u * sy + line) * W + 4u * sx;
    px[pix >> 2] = v;
    if (clear) sq[t] = make_uint4(0u, 0u, 0u, 0u);
}

// ---------------------------------------------------------------------------
// vox_read_raw_data's sample loop.  A workgroup of 256 threads owns a tile
// of kTile consecutive samples, 16 per thread.  Pass 1 counts the selected
// samples of every tile, pass 2 turns the counts into tile offsets (one
// workgroup) and the total, pass 3 recounts, ranks the threads of the tile
// and writes the voxel corners in raster order.
// ---------------------------------------------------------------------------
constexpr uint32_t kPer = 16;
constexpr uint32_t kTile = 256 * kPer;

struct RawArgs {
    const uint8_t *data;
    uint64_t n;          // samples
    uint32_t d1, d2;     // dim[1], dim[2]
    uint64_t lo, hi;     // the test lo <= sample < hi
    float v0, v1, v2;    // vox_voxel
    int aligned;         // data is 16-byte aligned
};

// The kPer samples from sample `first` on (little endian, SS bytes each)
// and how many of them exist.
template <int SS>
__device__ __forceinline__ uint32_t load_samples(const RawArgs &a,
                                                 uint64_t first,
                                                 uint32_t v[kPer]) {
    const uint64_t left = a.n > first ? a.n - first : 0;
    const uint32_t m = left < kPer ? (uint32_t)left : kPer;
    if (m == kPer && a.aligned) {
        // kPer * SS bytes starting at a multiple of 16
        uint32_t w[4 * SS + 1];
        const uint4 *q = reinterpret_cast<const uint4 *>(a.data + first * SS);
#pragma unroll
        for (int k = 0; k < SS; k++) {
            const uint4 x = q[k];
            w[4 * k + 0] = x.x;
            w[4 * k + 1] = x.y;
            w[4 * k + 2] = x.z;
            w[4 * k + 3] = x.w;
        }
        w[4 * SS] = 0;
#pragma unroll
        for (uint32_t k = 0; k < kPer; k++) {
            const uint32_t byte = k * SS, wi = byte >> 2, sh = 8u * (byte & 3u);
            const uint64_t pair = ((uint64_t)w[wi + 1] << 32) | w[wi];
            const uint32_t x = (uint32_t)(pair >> sh);
            v[k] = SS == 4 ? x : (x & ((1u << (8 * SS)) - 1u));
        }
    } else {
        for (uint32_t k = 0; k < kPer; k++) {
            uint32_t x = 0;
            if (k < m) {
                const uint8_t *p = a.data + (first + k) * SS;
#pragma unroll
                for (int l = 0; l < SS; l++) x |= (uint32_t)p[l] << (8 * l);
            }
            v[k] = x;
        }
    }
    return m;
}

__device__ __forceinline__ bool selected(const RawArgs &a, uint32_t v) {
    return (uint64_t)v >= a.lo && (uint64_t)v < a.hi;
}

// Sum of x over the workgroup (256 threads); every thread gets the sum,
// and `below` the sum over the threads before it.
__device__ __forceinline__ uint32_t block_scan(uint32_t x, uint32_t &below) {
    __shared__ uint32_t wsum[4];
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    uint32_t incl = x;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(incl, off);
        if (lane >= (uint32_t)off) incl += y;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (uint32_t i = 0; i < 4; i++) {
        before += i < w ? wsum[i] : 0u;
        total += wsum[i];
    }
    __syncthreads();
    below = before + incl - x;
    return total;
}

template <int SS>
__global__ void __launch_bounds__(256)
raw_count_kernel(RawArgs a, uint32_t *__restrict__ tile_count) {
    const uint64_t first = (uint64_t)blockIdx.x * kTile + threadIdx.x * kPer;
    uint32_t v[kPer];
    const uint32_t m = load_samples<SS>(a, first, v);
    uint32_t c = 0;
#pragma unroll
    for (uint32_t k = 0; k < kPer; k++) c += (k < m && selected(a, v[k])) ? 1u : 0u;
    uint32_t below;
    const uint32_t total = block_scan(c, below);
    if (threadIdx.x == 0) tile_count[blockIdx.x] = total;
}

// Exclusive prefix sum of the tile counts, in place, by one workgroup;
// the total goes to *count.
__global__ void __launch_bounds__(256)
raw_scan_kernel(uint32_t *__restrict__ tile, uint32_t ntiles,
                unsigned long long *__restrict__ count) {
    uint32_t carry = 0;
    for (uint32_t base = 0; base < ntiles; base += 256) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t x = i < ntiles ? tile[i] : 0u;
        uint32_t below;
        const uint32_t total = block_scan(x, below);
        if (i < ntiles) tile[i] = carry + below;
        carry += total;
    }
    if (threadIdx.x == 0) *count = carry;
}

template <int SS>
__global__ void __launch_bounds__(256)
raw_write_kernel(RawArgs a, const uint32_t *__restrict__ tile_base,
                 float *__restrict__ xyz, uint64_t capacity) {
    const uint64_t first = (uint64_t)blockIdx.x * kTile + threadIdx.x * kPer;
    uint32_t v[kPer];
    const uint32_t m = load_samples<SS>(a, first, v);
    uint32_t c = 0;
#pragma unroll
    for (uint32_t k = 0; k < kPer; k++) c += (k < m && selected(a, v[k])) ? 1u : 0u;
    uint32_t below;
    block_scan(c, below);
    if (c == 0) return;
    uint64_t out = (uint64_t)tile_base[blockIdx.x] + below;
    // raster position of the first sample: first = (i * d1 + j) * d2 + k
    const uint64_t ij = first / a.d2;
    uint32_t k = (uint32_t)(first - ij * a.d2);
    uint32_t j = (uint32_t)(ij % a.d1);
    uint32_t i = (uint32_t)(ij / a.d1);
#pragma unroll
    for (uint32_t s = 0; s < kPer; s++) {
        if (s < m && selected(a, v[s])) {
            if (out < capacity) {
                // datareader.c:66-68: i*vox_voxel[0] etc. in float
                xyz[3 * out + 0] = (float)i * a.v0;
                xyz[3 * out + 1] = (float)j * a.v1;
                xyz[3 * out + 2] = (float)k * a.v2;
            }
            out++;
        }
        if (++k == a.d2) {
            k = 0;
            if (++j == a.d1) {
                j = 0;
                i++;
            }
        }
    }
}

template <int SS>
int raw_select_ss(const RawArgs &a, uint32_t ntiles, uint32_t *tiles,
                  unsigned long long *count, float *xyz, uint64_t capacity,
                  hipStream_t stream) {
    hipLaunchKernelGGL(raw_count_kernel<SS>, dim3(ntiles), dim3(256), 0,
                       stream, a, tiles);
    VOXG_HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(raw_scan_kernel, dim3(1), dim3(256), 0, stream, tiles,
                       ntiles, count);
    VOXG_HIPCHK(hipGetLastError());
    if (xyz) {
        hipLaunchKernelGGL(raw_write_kernel<SS>, dim3(ntiles), dim3(256), 0,
                           stream, a, tiles, xyz, capacity);
        VOXG_HIPCHK(hipGetLastError());
    }
    return 0;
}

}  // namespace

extern "C" {

int voxg_redraw(uint32_t *squares, uint32_t *pixels, size_t n_frames, int W,
                int H, int clear, hipStream_t stream) {
    if (W <= 0 || H <= 0 || (W & 0xf) || (H & 0x3))
        return fail(VOXG_EINVAL, "W must be a positive multiple of 16 and H of 4");
    if (n_frames == 0) return 0;
    if (!squares || !pixels) return fail(VOXG_EINVAL, "buffer is NULL");
    if (((uintptr_t)squares | (uintptr_t)pixels) & 15u)
        return fail(VOXG_EINVAL, "squares and pixels must be 16-byte aligned");
    const uint64_t per = (uint64_t)W * (uint64_t)H;
    if (n_frames > (1ull << 40) / per)
        return fail(VOXG_ERANGE, "n_frames * W * H too large");
    const uint64_t nthreads = n_frames * per / 4u;
    const uint32_t ws = (uint32_t)W / 4u, bpf = ws * ((uint32_t)H / 4u);
    const uint64_t grid = (nthreads + 255) / 256;
    if (grid > 0x7fffffffull) return fail(VOXG_ERANGE, "frames too large");
    hipLaunchKernelGGL(redraw_kernel, dim3((uint32_t)grid), dim3(256), 0,
                       stream, reinterpret_cast<uint4 *>(squares),
                       reinterpret_cast<uint4 *>(pixels), ws, bpf,
                       (uint32_t)W, nthreads, clear);
    VOXG_HIPCHK(hipGetLastError());
    return 0;
}

// dim[0] * dim[1] * dim[2] into *n, or false when that reaches `limit`:
// tested factor by factor, so that no product wraps around
static bool volume_samples(const unsigned dim[3], uint64_t limit,
                           uint64_t *n) {
    *n = 0;
    if (dim[0] == 0 || dim[1] == 0 || dim[2] == 0) return true;
    uint64_t p = 1;
    for (int i = 0; i < 3; i++) {
        if (p > (limit - 1) / dim[i]) return false;
        p *= dim[i];
    }
    *n = p;
    return true;
}

int voxg_raw_select(const void *samples, const unsigned dim[3],
                    unsigned samplesize, uint64_t lo, uint64_t hi,
                    const float voxel[3], float *xyz_out, size_t capacity,
                    unsigned long long *count_out, hipStream_t stream) {
    if (!dim || !count_out) return fail(VOXG_EINVAL, "dim or count_out is NULL");
    if (samplesize < 1 || samplesize > 4)
        return fail(VOXG_EINVAL, "samplesize must be 1, 2, 3 or 4");
    uint64_t n = 0;
    if (!volume_samples(dim, (1ull << 32) - kTile, &n))
        return fail(VOXG_ERANGE, "volume must hold fewer than 2^32 samples");
    if (n && !samples) return fail(VOXG_EINVAL, "samples is NULL");
    const uint32_t ntiles = (uint32_t)((n + kTile - 1) / kTile);
    void *scr = nullptr;
    int rc = voxg_internal::scratch(stream, (size_t)(ntiles + 1) * 4, &scr);
    if (rc) return rc;
    RawArgs a;
    a.data = static_cast<const uint8_t *>(samples);
    a.n = n;
    a.d1 = dim[1];
    a.d2 = dim[2];
    a.lo = lo;
    a.hi = hi;
    a.v0 = voxel ? voxel[0] : 1.0f;
    a.v1 = voxel ? voxel[1] : 1.0f;
    a.v2 = voxel ? voxel[2] : 1.0f;
    a.aligned = ((uintptr_t)samples & 15u) == 0;
    if (n == 0) {
        VOXG_HIPCHK(hipMemsetAsync(count_out, 0, sizeof(*count_out), stream));
        return 0;
    }
    uint32_t *tiles = static_cast<uint32_t *>(scr);
    switch (samplesize) {
    case 1: return raw_select_ss<1>(a, ntiles, tiles, count_out, xyz_out, capacity, stream);
    case 2: return raw_select_ss<2>(a, ntiles, tiles, count_out, xyz_out, capacity, stream);
    case 3: return raw_select_ss<3>(a, ntiles, tiles, count_out, xyz_out, capacity, stream);
    default: return raw_select_ss<4>(a, ntiles, tiles, count_out, xyz_out, capacity, stream);
    }
}

int voxg_read_raw_data(const char *filename, const unsigned dim[3],
                       unsigned samplesize, uint64_t lo, uint64_t hi,
                       const float voxel[3], voxg_tree **out) {
    if (!out) return fail(VOXG_EINVAL, "out is NULL");
    *out = nullptr;
    if (!filename || !dim) return fail(VOXG_EINVAL, "filename or dim is NULL");
    if (samplesize < 1 || samplesize > 4)
        return fail(VOXG_EINVAL, "samplesize must be 1, 2, 3 or 4");
    uint64_t n = 0;
    if (!volume_samples(dim, (1ull << 32) - kTile, &n))
        return fail(VOXG_ERANGE, "volume must hold fewer than 2^32 samples");
    const int fd = open(filename, O_RDONLY);
    if (fd < 0) return fail(VOXG_EIO, std::string(filename) + ": " + strerror(errno));
    struct stat sb;
    if (fstat(fd, &sb) != 0) {
        int e = errno;
        close(fd);
        return fail(VOXG_EIO, std::string(filename) + ": " + strerror(e));
    }
    // datareader.c:39-42
    if ((uint64_t)sb.st_size != n * samplesize) {
        close(fd);
        return fail(VOXG_EINVAL, "Wrong size of dataset");
    }
    const size_t bytes = (size_t)(n * samplesize);
    uint8_t *host = nullptr, *dev = nullptr;
    unsigned long long *dcount = nullptr;
    float *dxyz = nullptr, *hxyz = nullptr;
    hipStream_t s = nullptr;
    int rc = 0;
    auto cleanup = [&]() {  // best effort: the first error is reported
        if (s) {
            voxg_internal::release_stream(s);
            (void)hipStreamDestroy(s);
        }
        if (host) (void)hipHostFree(host);
        if (hxyz) (void)hipHostFree(hxyz);
        if (dev) (void)hipFree(dev);
        if (dxyz) (void)hipFree(dxyz);
        if (dcount) (void)hipFree(dcount);
        close(fd);
    };
#define VOXG_TRY(expr)                                                     \
    do {                                                                   \
        hipError_t e_ = (expr);                                            \
        if (e_ != hipSuccess) {                                            \
            rc = fail(VOXG_EIO, std::string(#expr ": ") +                  \
                                    hipGetErrorString(e_));                \
            cleanup();                                                     \
            return rc;                                                     \
        }                                                                  \
    } while (0)
    VOXG_TRY(hipStreamCreate(&s));
    VOXG_TRY(hipHostMalloc((void **)&host, bytes ? bytes : 1));
    VOXG_TRY(hipMalloc((void **)&dev, bytes ? bytes : 16));
    VOXG_TRY(hipMalloc((void **)&dcount, sizeof(unsigned long long)));
    size_t got = 0;
    while (got < bytes) {
        ssize_t r = read(fd, host + got, bytes - got);
        if (r <= 0) {
            rc = fail(VOXG_EIO, std::string(filename) + ": " +
                                    (r < 0 ? strerror(errno) : "short read"));
            cleanup();
            return rc;
        }
        got += (size_t)r;
    }
    VOXG_TRY(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, s));
    unsigned long long total = 0;
    rc = voxg_raw_select(dev, dim, samplesize, lo, hi, voxel, nullptr, 0,
                         dcount, s);
    if (rc) {
        cleanup();
        return rc;
    }
    VOXG_TRY(hipMemcpyAsync(&total, dcount, sizeof(total),
                            hipMemcpyDeviceToHost, s));
    VOXG_TRY(hipStreamSynchronize(s));
    if (total) {
        VOXG_TRY(hipMalloc((void **)&dxyz, (size_t)total * 3 * sizeof(float)));
        VOXG_TRY(hipHostMalloc((void **)&hxyz, (size_t)total * 3 * sizeof(float)));
        rc = voxg_raw_select(dev, dim, samplesize, lo, hi, voxel, dxyz, total,
                             dcount, s);
        if (rc) {
            cleanup();
            return rc;
        }
        VOXG_TRY(hipMemcpyAsync(hxyz, dxyz, (size_t)total * 3 * sizeof(float),
                                hipMemcpyDeviceToHost, s));
        VOXG_TRY(hipStreamSynchronize(s));
    }
#undef VOXG_TRY
    rc = voxg_tree_build(hxyz, (size_t)total, 3, voxel, out);
    cleanup();
    return rc;
}

}  // extern "C"
