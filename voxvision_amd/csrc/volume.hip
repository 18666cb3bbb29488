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
    // streamed once: non-temporal loads and stores (4.9 -> 6.0 TB/s on the
    // benchmark frames)
    typedef unsigned int u4 __attribute__((ext_vector_type(4)));
    const u4 v = __builtin_nontemporal_load(reinterpret_cast<const u4 *>(sq + t));
    const uint64_t pix = f * (uint64_t)bpf * 16u +
                         (uint64_t)(4u * sy + line) * W + 4u * sx;
    __builtin_nontemporal_store(v, reinterpret_cast<u4 *>(px + (pix >> 2)));
    if (clear) __builtin_nontemporal_store(u4{0u, 0u, 0u, 0u}, reinterpret_cast<u4 *>(sq + t));
}

// ---------------------------------------------------------------------------
// vox_read_raw_data's sample loop as an ordered compaction.  A workgroup of
// 256 threads owns a tile of kTile consecutive samples, kPer per thread.
// Phase 1: raw_count_kernel reads the samples once (32 to 128 bytes per
// thread, enough loads in flight to stream the volume), stores each
// thread's selection as a 32-bit mask and counts the selected samples of
// every tile; raw_scan_kernel turns the counts into offsets within groups
// of kGroup tiles (many workgroups, coalesced), raw_total_kernel the group
// sums into group offsets and the total.  Phase 2: raw_write_kernel reads
// only the masks, ranks the threads of a tile, stages the tile's voxel
// corners in LDS in raster order and stores them with consecutive threads
// on consecutive words.  The masks and offsets live in a workspace (the
// caller's, or the stream's scratch), so that the two phases can be split
// around a host allocation of the exact output.
// ---------------------------------------------------------------------------
constexpr uint32_t kPer = 32;
constexpr uint32_t kTile = 256 * kPer;
constexpr uint32_t kGroup = 1024;   // tiles per scan group
constexpr uint32_t kTotalThreads = 1024;
constexpr uint32_t kStage = 1024;   // corners staged in LDS per round
// fewer than 2^32 samples: at most 2^19 tiles, 512 groups
static_assert((1ull << 32) / kTile / kGroup <= kTotalThreads,
              "one workgroup scans the group sums");

struct RawArgs {
    const uint8_t *data;
    uint64_t n;          // samples
    uint64_t lo, hi;     // the test lo <= sample < hi
    int aligned;         // data is 16-byte aligned
};

// Workspace of a volume of n samples: the tile offsets (ntiles + 1 words,
// the last one the total), the group offsets, then one 32-bit mask per
// thread, each part padded to 256 bytes.
struct RawWork {
    uint32_t *tiles, *groups, *masks;
};

uint32_t raw_tiles(uint64_t n) { return (uint32_t)((n + kTile - 1) / kTile); }
uint32_t raw_groups(uint64_t n) { return (raw_tiles(n) + kGroup - 1) / kGroup; }

size_t pad256(size_t b) { return (b + 255) / 256 * 256; }

size_t raw_work_bytes(uint64_t n) {
    return pad256(((size_t)raw_tiles(n) + 1) * 4) +
           pad256((size_t)raw_groups(n) * 4) + (size_t)raw_tiles(n) * 256 * 4;
}

RawWork raw_work(void *base, uint64_t n) {
    uint8_t *p = static_cast<uint8_t *>(base);
    RawWork w;
    w.tiles = reinterpret_cast<uint32_t *>(p);
    p += pad256(((size_t)raw_tiles(n) + 1) * 4);
    w.groups = reinterpret_cast<uint32_t *>(p);
    p += pad256((size_t)raw_groups(n) * 4);
    w.masks = reinterpret_cast<uint32_t *>(p);
    return w;
}

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
        constexpr int NQ = kPer * SS / 16;
        uint32_t w[4 * NQ + 1];
        const uint4 *q = reinterpret_cast<const uint4 *>(a.data + first * SS);
#pragma unroll
        for (int k = 0; k < NQ; k++) {
            typedef unsigned int u4 __attribute__((ext_vector_type(4)));
            const u4 xv = __builtin_nontemporal_load(reinterpret_cast<const u4 *>(q + k));
            const uint4 x = make_uint4(xv.x, xv.y, xv.z, xv.w);
            w[4 * k + 0] = x.x;
            w[4 * k + 1] = x.y;
            w[4 * k + 2] = x.z;
            w[4 * k + 3] = x.w;
        }
        w[4 * NQ] = 0;
#pragma unroll
        for (uint32_t k = 0; k < kPer; k++) {
            const uint32_t byte = k * SS, wi = byte >> 2, sh = 8u * (byte & 3u);
            const uint64_t pair = ((uint64_t)w[wi + 1] << 32) | w[wi];
            const uint32_t x = (uint32_t)(pair >> sh);
            // the sample's SS low bytes (all four when SS == 4)
            v[k] = x & (0xffffffffu >> (32 - 8 * SS));
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

// Sum of x over the workgroup (NW waves); every thread gets the sum, and
// `below` the sum over the threads before it.
template <uint32_t NW>
__device__ __forceinline__ uint32_t block_scan(uint32_t x, uint32_t &below) {
    __shared__ uint32_t wsum[NW];
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
    for (uint32_t i = 0; i < NW; i++) {
        before += i < w ? wsum[i] : 0u;
        total += wsum[i];
    }
    __syncthreads();
    below = before + incl - x;
    return total;
}

template <int SS>
__global__ void __launch_bounds__(256)
raw_count_kernel(RawArgs a, RawWork w) {
    const uint64_t g = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    uint32_t v[kPer];
    const uint32_t m = load_samples<SS>(a, g * kPer, v);
    uint32_t mask = 0;
#pragma unroll
    for (uint32_t k = 0; k < kPer; k++)
        mask |= (k < m && selected(a, v[k])) ? 1u << k : 0u;
    w.masks[g] = mask;
    uint32_t below;
    const uint32_t total = block_scan<4>((uint32_t)__popc(mask), below);
    if (threadIdx.x == 0) w.tiles[blockIdx.x] = total;
}

// Offsets of the tiles within their group of kGroup, in place, and the
// group's sum.
__global__ void __launch_bounds__(kGroup)
raw_scan_kernel(RawWork w, uint32_t ntiles) {
    const uint32_t t = blockIdx.x * kGroup + threadIdx.x;
    const uint32_t x = t < ntiles ? w.tiles[t] : 0u;
    uint32_t below;
    const uint32_t total = block_scan<kGroup / 64>(x, below);
    if (t < ntiles) w.tiles[t] = below;
    if (threadIdx.x == 0) w.groups[blockIdx.x] = total;
}

// Offsets of the groups, in place; the total after the last tile offset
// and in *count.
__global__ void __launch_bounds__(kTotalThreads)
raw_total_kernel(RawWork w, uint32_t ntiles, uint32_t ngroups,
                 unsigned long long *__restrict__ count) {
    const uint32_t x = threadIdx.x < ngroups ? w.groups[threadIdx.x] : 0u;
    uint32_t below;
    const uint32_t total = block_scan<kTotalThreads / 64>(x, below);
    if (threadIdx.x < ngroups) w.groups[threadIdx.x] = below;
    if (threadIdx.x == 0) {
        w.tiles[ntiles] = total;
        if (count) *count = total;
    }
}

struct RawOut {
    uint64_t n;
    uint32_t d1, d2;
    float v0, v1, v2;
};

__global__ void __launch_bounds__(256)
raw_write_kernel(RawOut a, RawWork w, float *__restrict__ xyz,
                 uint64_t capacity) {
    __shared__ float stage[3 * kStage];
    const uint64_t g = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    const uint64_t first = g * kPer;
    const uint32_t mask = first < a.n ? w.masks[g] : 0u;
    uint32_t below;
    const uint32_t total = block_scan<4>((uint32_t)__popc(mask), below);
    // the tile's first corner (the same for every thread of the tile)
    const uint64_t base = (uint64_t)w.tiles[blockIdx.x] +
                          w.groups[blockIdx.x / kGroup];
    if (base >= capacity) return;
    // raster position of the first sample: first = (i * d1 + j) * d2 + k
    const uint64_t ij = first / a.d2;
    const uint32_t k0 = (uint32_t)(first - ij * a.d2);
    const uint32_t j0 = (uint32_t)(ij % a.d1), i0 = (uint32_t)(ij / a.d1);
    for (uint32_t r0 = 0; r0 < total; r0 += kStage) {
        // the corners of ranks r0 .. r0 + kStage - 1 into LDS
        if (mask && below < r0 + kStage && below + __popc(mask) > r0) {
            // the set bits in order; the raster position advances by the
            // gap to each (at most kPer steps in all)
            uint32_t rank = below, k = k0, j = j0, i = i0, at = 0, mm = mask;
            while (mm) {
                const uint32_t s = (uint32_t)__ffs(mm) - 1u;
                mm &= mm - 1u;
                k += s - at;
                at = s;
                while (k >= a.d2) {
                    k -= a.d2;
                    if (++j == a.d1) {
                        j = 0;
                        i++;
                    }
                }
                if (rank - r0 < kStage) {
                    // datareader.c:66-68: i*vox_voxel[0] etc. in float
                    float *c = stage + 3u * (rank - r0);
                    c[0] = (float)i * a.v0;
                    c[1] = (float)j * a.v1;
                    c[2] = (float)k * a.v2;
                }
                rank++;
            }
        }
        __syncthreads();
        const uint32_t cnt = min(kStage, total - r0);
        const uint64_t o0 = base + r0;
        for (uint32_t t = threadIdx.x; t < 3u * cnt; t += 256u)
            if (o0 + t / 3u < capacity)
                __builtin_nontemporal_store(stage[t], xyz + 3u * o0 + t);
        __syncthreads();
    }
}

template <int SS>
void raw_count_ss(const RawArgs &a, uint32_t ntiles, const RawWork &w,
                  hipStream_t stream) {
    hipLaunchKernelGGL(raw_count_kernel<SS>, dim3(ntiles), dim3(256), 0,
                       stream, a, w);
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

// Argument checks shared by the raw entry points: the sample count, or an
// error.
static int raw_samples(const unsigned dim[3], uint64_t *n) {
    if (!dim) return fail(VOXG_EINVAL, "dim is NULL");
    if (!volume_samples(dim, (1ull << 32) - kTile, n))
        return fail(VOXG_ERANGE, "volume must hold fewer than 2^32 samples");
    return 0;
}

size_t voxg_raw_workspace_bytes(const unsigned dim[3]) {
    uint64_t n = 0;
    if (!dim || !volume_samples(dim, (1ull << 32) - kTile, &n)) return 0;
    return raw_work_bytes(n);
}

int voxg_raw_count(const void *samples, const unsigned dim[3],
                   unsigned samplesize, uint64_t lo, uint64_t hi, void *work,
                   unsigned long long *count_out, hipStream_t stream) {
    uint64_t n = 0;
    int rc = raw_samples(dim, &n);
    if (rc) return rc;
    if (samplesize < 1 || samplesize > 4)
        return fail(VOXG_EINVAL, "samplesize must be 1, 2, 3 or 4");
    if (!work) return fail(VOXG_EINVAL, "work is NULL");
    if (n && !samples) return fail(VOXG_EINVAL, "samples is NULL");
    if ((uintptr_t)work & 15u)
        return fail(VOXG_EINVAL, "work must be 16-byte aligned");
    const RawWork w = raw_work(work, n);
    const uint32_t ntiles = raw_tiles(n), ngroups = raw_groups(n);
    if (ntiles) {
        RawArgs a;
        a.data = static_cast<const uint8_t *>(samples);
        a.n = n;
        a.lo = lo;
        a.hi = hi;
        a.aligned = ((uintptr_t)samples & 15u) == 0;
        switch (samplesize) {
        case 1: raw_count_ss<1>(a, ntiles, w, stream); break;
        case 2: raw_count_ss<2>(a, ntiles, w, stream); break;
        case 3: raw_count_ss<3>(a, ntiles, w, stream); break;
        default: raw_count_ss<4>(a, ntiles, w, stream); break;
        }
        VOXG_HIPCHK(hipGetLastError());
        hipLaunchKernelGGL(raw_scan_kernel, dim3(ngroups), dim3(kGroup), 0,
                           stream, w, ntiles);
        VOXG_HIPCHK(hipGetLastError());
    }
    // the group offsets and the total (zero with no tile)
    hipLaunchKernelGGL(raw_total_kernel, dim3(1), dim3(kTotalThreads), 0,
                       stream, w, ntiles, ngroups, count_out);
    VOXG_HIPCHK(hipGetLastError());
    return 0;
}

int voxg_raw_write(const unsigned dim[3], const float voxel[3],
                   const void *work, float *xyz_out, size_t capacity,
                   hipStream_t stream) {
    uint64_t n = 0;
    int rc = raw_samples(dim, &n);
    if (rc) return rc;
    if (!work) return fail(VOXG_EINVAL, "work is NULL");
    if ((uintptr_t)work & 15u)
        return fail(VOXG_EINVAL, "work must be 16-byte aligned");
    const uint32_t ntiles = raw_tiles(n);
    if (!ntiles || capacity == 0) return 0;
    if (!xyz_out) return fail(VOXG_EINVAL, "xyz_out is NULL");
    RawOut a;
    a.n = n;
    a.d1 = dim[1];
    a.d2 = dim[2];
    a.v0 = voxel ? voxel[0] : 1.0f;
    a.v1 = voxel ? voxel[1] : 1.0f;
    a.v2 = voxel ? voxel[2] : 1.0f;
    hipLaunchKernelGGL(raw_write_kernel, dim3(ntiles), dim3(256), 0, stream,
                       a, raw_work(const_cast<void *>(work), n), xyz_out,
                       (uint64_t)capacity);
    VOXG_HIPCHK(hipGetLastError());
    return 0;
}

int voxg_raw_select(const void *samples, const unsigned dim[3],
                    unsigned samplesize, uint64_t lo, uint64_t hi,
                    const float voxel[3], float *xyz_out, size_t capacity,
                    unsigned long long *count_out, hipStream_t stream) {
    if (!count_out) return fail(VOXG_EINVAL, "count_out is NULL");
    uint64_t n = 0;
    int rc = raw_samples(dim, &n);
    if (rc) return rc;
    void *scr = nullptr;
    rc = voxg_internal::scratch(stream, raw_work_bytes(n), &scr);
    if (rc) return rc;
    rc = voxg_raw_count(samples, dim, samplesize, lo, hi, scr, count_out,
                        stream);
    if (rc || !xyz_out) return rc;
    return voxg_raw_write(dim, voxel, scr, xyz_out, capacity, stream);
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
    void *work = nullptr;
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
        if (work) (void)hipFree(work);
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
    VOXG_TRY(hipMalloc(&work, raw_work_bytes(n)));
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
    rc = voxg_raw_count(dev, dim, samplesize, lo, hi, work, dcount, s);
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
        rc = voxg_raw_write(dim, voxel, work, dxyz, total, s);
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
