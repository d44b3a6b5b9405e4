// tensorio.hip -- the on-disk quantized tensor format (.tdxq) and its writer and loader.
//
// The format stores a dense column-major tensor in one of the library's storages (FP32, BF16,
// or FP8-E4M3 with one FP32 scale per mode-1 slice X(i,:,...)) exactly as the MTTKRP kernels
// read it, so a loaded tensor goes to tdx_mttkrp_ex / the CP-ALS driver without conversion.
// It extends the reference's precision-reduction step f2h (src/float2half.cu:14-18), which
// only converts in memory.  Layout (all integers little-endian):
//
//   offset  size  field
//        0     4  magic "TDXQ"
//        4     4  version (1)
//        8     4  order (2..4)
//       12     4  storage: 0 FP32, 1 BF16 (bit patterns, RNE), 2 FP8 E4M3FN (OCP) with scales
//       16    32  dims[4] (int64; 1 past the order)
//       48     8  n_scales (int64): dims[0] for FP8 storage, 0 otherwise
//       56     8  data_bytes (int64): prod(dims) * element size
//       64     8  FNV-1a 64 of bytes 0..63 (header check)
//       72    56  zero
//      128        scales: n_scales FP32 (value = max |X(i,:,...)| / 448, 1 for an all-zero slice),
//                 zero padded to a multiple of 128 bytes
//   128 + S       data: data_bytes, column-major X[i + j I + k I J (+ l I J K)]
//
// Host pointers are read and written directly; device pointers go through two pinned staging
// buffers so the file I/O of one chunk overlaps the PCIe copy of the other.
#include "common.hpp"
#include <stdio.h>
#include <string.h>

namespace tdx {

static constexpr uint32_t kTdxqVersion = 1;
static constexpr size_t kHeaderBytes = 128;
static constexpr size_t kChunk = size_t(64) << 20;

static uint64_t fnv1a(const unsigned char *p, size_t n) {
    uint64_t h = 1469598103934665603ull;
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

static size_t elem_bytes(int storage) { return storage == TDX_F32 ? 4 : storage == TDX_BF16 ? 2 : 1; }
static size_t pad128(size_t b) { return (b + 127) & ~size_t(127); }

static void put32(unsigned char *h, size_t off, uint32_t v) { memcpy(h + off, &v, 4); }
static void put64(unsigned char *h, size_t off, int64_t v) { memcpy(h + off, &v, 8); }
static uint32_t get32(const unsigned char *h, size_t off) { uint32_t v; memcpy(&v, h + off, 4); return v; }
static int64_t get64(const unsigned char *h, size_t off) { int64_t v; memcpy(&v, h + off, 8); return v; }

static int parse_header(const unsigned char *h, tdx_tensor_info *info) {
    if (memcmp(h, "TDXQ", 4) != 0 || get32(h, 4) != kTdxqVersion) return TDX_ERR_IO;
    if ((uint64_t)get64(h, 64) != fnv1a(h, 64)) return TDX_ERR_IO;
    tdx_tensor_info in{};
    in.order = (int)get32(h, 8);
    in.storage = (int)get32(h, 12);
    if (in.order < 2 || in.order > 4) return TDX_ERR_IO;
    if (in.storage != TDX_F32 && in.storage != TDX_BF16 && in.storage != TDX_FP8_E4M3) return TDX_ERR_IO;
    int64_t n = 1;
    for (int m = 0; m < 4; ++m) {
        in.dims[m] = get64(h, 16 + 8 * m);
        if (in.dims[m] <= 0 || (m >= in.order && in.dims[m] != 1)) return TDX_ERR_IO;
        n *= in.dims[m];
    }
    in.n_scales = get64(h, 48);
    in.data_bytes = get64(h, 56);
    if (in.n_scales != (in.storage == TDX_FP8_E4M3 ? in.dims[0] : 0)) return TDX_ERR_IO;
    if (in.data_bytes != n * (int64_t)elem_bytes(in.storage)) return TDX_ERR_IO;
    *info = in;
    return TDX_OK;
}

// Pinned double buffer for device transfers; frees on scope exit.
struct Staging {
    void *buf[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    ~Staging() {
        for (int b = 0; b < 2; ++b) {
            if (done[b]) {
                (void)hipEventSynchronize(done[b]);
                (void)hipEventDestroy(done[b]);
            }
            if (buf[b]) (void)hipHostFree(buf[b]);
        }
    }
    int init(size_t bytes) {
        for (int b = 0; b < 2; ++b) {
            TDX_HIP(hipHostMalloc(&buf[b], bytes, hipHostMallocDefault));
            TDX_HIP(hipEventCreateWithFlags(&done[b], hipEventDisableTiming));
        }
        return TDX_OK;
    }
};

// file -> (host or device) memory, in chunks
static int read_into(FILE *f, void *dst, size_t bytes, int on_device, Staging *stg) {
    if (!on_device) return fread(dst, 1, bytes, f) == bytes ? TDX_OK : TDX_ERR_IO;
    hipStream_t s = stream();
    size_t off = 0;
    int b = 0;
    while (off < bytes) {
        const size_t n = bytes - off < kChunk ? bytes - off : kChunk;
        TDX_HIP(hipEventSynchronize(stg->done[b]));  // the copy that used this buffer has finished
        if (fread(stg->buf[b], 1, n, f) != n) return TDX_ERR_IO;
        TDX_HIP(hipMemcpyAsync((char *)dst + off, stg->buf[b], n, hipMemcpyHostToDevice, s));
        TDX_HIP(hipEventRecord(stg->done[b], s));
        off += n;
        b ^= 1;
    }
    TDX_HIP(hipStreamSynchronize(s));
    return TDX_OK;
}

// (host or device) memory -> file, in chunks
static int write_from(FILE *f, const void *src, size_t bytes, int on_device, Staging *stg) {
    if (!on_device) return fwrite(src, 1, bytes, f) == bytes ? TDX_OK : TDX_ERR_IO;
    hipStream_t s = stream();
    size_t off = 0, pend = 0;
    int b = 0;
    // copy chunk i into one buffer while chunk i-1 is written from the other
    while (off < bytes || pend) {
        size_t n = 0;
        if (off < bytes) {
            n = bytes - off < kChunk ? bytes - off : kChunk;
            TDX_HIP(hipMemcpyAsync(stg->buf[b], (const char *)src + off, n, hipMemcpyDeviceToHost, s));
            TDX_HIP(hipEventRecord(stg->done[b], s));
        }
        if (pend) {
            TDX_HIP(hipEventSynchronize(stg->done[b ^ 1]));
            if (fwrite(stg->buf[b ^ 1], 1, pend, f) != pend) return TDX_ERR_IO;
        }
        pend = n;
        off += n;
        b ^= 1;
    }
    return TDX_OK;
}

struct FileGuard {
    FILE *f;
    ~FileGuard() {
        if (f) fclose(f);
    }
};

}  // namespace tdx

using namespace tdx;

extern "C" {

int tdx_tensor_info_read(const char *path, tdx_tensor_info *info) {
    TDX_CHECK_ARG(path && info);
    FileGuard g{fopen(path, "rb")};
    if (!g.f) return record(TDX_ERR_IO);
    unsigned char h[kHeaderBytes];
    if (fread(h, 1, kHeaderBytes, g.f) != kHeaderBytes) return record(TDX_ERR_IO);
    return record(parse_header(h, info));
}

int tdx_tensor_save(const char *path, const void *X, int storage, const float *scales, int order, const long *dims,
                    int on_device) {
    TDX_CHECK_ARG(path && X && dims && order >= 2 && order <= 4);
    TDX_CHECK_ARG(storage == TDX_F32 || storage == TDX_BF16 || storage == TDX_FP8_E4M3);
    TDX_CHECK_ARG(storage != TDX_FP8_E4M3 || scales);
    unsigned char h[kHeaderBytes];
    memset(h, 0, sizeof(h));
    memcpy(h, "TDXQ", 4);
    put32(h, 4, kTdxqVersion);
    put32(h, 8, (uint32_t)order);
    put32(h, 12, (uint32_t)storage);
    int64_t n = 1;
    for (int m = 0; m < 4; ++m) {
        const int64_t d = m < order ? dims[m] : 1;
        TDX_CHECK_ARG(d > 0);
        put64(h, 16 + 8 * m, d);
        n *= d;
    }
    const int64_t ns = storage == TDX_FP8_E4M3 ? dims[0] : 0;
    const size_t data_bytes = (size_t)n * elem_bytes(storage);
    put64(h, 48, ns);
    put64(h, 56, (int64_t)data_bytes);
    put64(h, 64, (int64_t)fnv1a(h, 64));
    Staging stg;
    if (on_device) {
        int st = stg.init(kChunk);
        if (st) return st;
    }
    FileGuard g{fopen(path, "wb")};
    if (!g.f) return record(TDX_ERR_IO);
    if (fwrite(h, 1, kHeaderBytes, g.f) != kHeaderBytes) return record(TDX_ERR_IO);
    if (ns) {
        const size_t sb = sizeof(float) * (size_t)ns, sp = pad128(sb);
        int st = write_from(g.f, scales, sb, on_device, &stg);
        if (st) return record(st);
        static const unsigned char zeros[128] = {0};
        if (sp > sb && fwrite(zeros, 1, sp - sb, g.f) != sp - sb) return record(TDX_ERR_IO);
    }
    int st = write_from(g.f, X, data_bytes, on_device, &stg);
    if (st) return record(st);
    if (fflush(g.f) != 0) return record(TDX_ERR_IO);
    return record(TDX_OK);
}

int tdx_tensor_load(const char *path, void *X, size_t x_capacity, float *scales, size_t scales_capacity,
                    int on_device, tdx_tensor_info *info) {
    TDX_CHECK_ARG(path && X);
    FileGuard g{fopen(path, "rb")};
    if (!g.f) return record(TDX_ERR_IO);
    unsigned char h[kHeaderBytes];
    if (fread(h, 1, kHeaderBytes, g.f) != kHeaderBytes) return record(TDX_ERR_IO);
    tdx_tensor_info in;
    int st = parse_header(h, &in);
    if (st) return record(st);
    if ((size_t)in.data_bytes > x_capacity) return record(TDX_ERR_ARG);
    if (in.n_scales && (!scales || sizeof(float) * (size_t)in.n_scales > scales_capacity)) return record(TDX_ERR_ARG);
    Staging stg;
    if (on_device) {
        st = stg.init(kChunk);
        if (st) return st;
    }
    if (in.n_scales) {
        const size_t sb = sizeof(float) * (size_t)in.n_scales, sp = pad128(sb);
        st = read_into(g.f, scales, sb, on_device, &stg);
        if (st) return record(st);
        if (sp > sb && fseek(g.f, (long)(sp - sb), SEEK_CUR) != 0) return record(TDX_ERR_IO);
    }
    st = read_into(g.f, X, (size_t)in.data_bytes, on_device, &stg);
    if (st) return record(st);
    if (info) *info = in;
    return record(TDX_OK);
}

}  // extern "C"
