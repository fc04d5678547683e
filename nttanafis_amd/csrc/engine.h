// Internal interface of the engine's host side, shared by the product and
// convolution pipelines (engine.hip) and the request server (server.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/nttanafis.h"
#include "arith.h"

namespace nttamd {

void set_error(const std::string& e);
void clear_error();

// Names the kernel launches of a scope for per-launch timing.
struct Timed {
    const char* saved;
    Timed(const char* n, hipStream_t s);
    ~Timed();
};

// ---------------------------------------------------------------------------
// Geometry of one device operation.

struct Geometry {
    uint32_t primes = 0;
    uint32_t width = 0;         // bits between operand coefficients
    uint32_t value_bits_a = 0;  // significant bits per coefficient
    uint32_t value_bits_b = 0;
    uint64_t digits_a = 0, digits_b = 0;
    uint64_t conv_len = 0;
    // Transform row: 2^log_t points (tail 0), or a composite row of
    // N + tail*N/4 points with N = 2^log_t, evaluated modulo x^N - 1 and the
    // factors of x^N + 1 at table nodes 4 (tail 1), 2 (tail 2), 2 and 6
    // (tail 3) (src/profile/planner.rs:565-591, src/transform/forward.rs:95-197).
    int log_t = 0;
    uint32_t tail = 0;
    // Unbalanced split of a product (src/profile/planner.rs:283-393): the
    // longer operand's coefficients run in `blocks` blocks of block_digits (a
    // multiple of 64), each on its own row of 2^log_t points against the
    // shorter operand's spectrum, transformed once; 0 for none.
    uint32_t blocks = 0;
    uint64_t block_digits = 0;
    uint32_t bound_bits = 0;  // bits of every reconstructed coefficient
    // Schoolbook product (no transform; the fields above still describe
    // the transform geometry the product would otherwise run).
    bool school = false;
    bool disjoint = false;
    bool is_signed = false;
    uint32_t stride = 0;  // limbs per result coefficient (convolution)
};

// Points of the transform row.
inline uint64_t row_len(const Geometry& g) {
    const uint64_t n = (uint64_t)1 << g.log_t;
    return n + g.tail * (n >> 2);
}

// One transform of a row: `log_n` points at table node `node`, `offset`
// points into the row; a composite row has two or three.
struct RowPart {
    int log_n;
    uint32_t node;
    uint64_t offset;
};
int row_parts(const Geometry& g, RowPart* parts);

// batched: geometry of a product inside a batch, which shares launches
// among its products and runs power-of-two rows only (whole or split).
bool select_product(uint64_t la, uint64_t lb, Geometry* out, bool batched = false);
bool conv_bound_bits(uint64_t terms, uint64_t max_a, uint64_t max_b, uint64_t* bits);
bool product_geometry(const ntt_mul_plan* plan, Geometry* g);
bool conv_geometry(const ntt_conv_plan* p, Geometry* g);
// ConvPlan::admits, src/convolve.rs:169-184.
bool conv_admits(const ntt_conv_plan* p, const uint64_t* a, size_t a_len, const uint64_t* b, size_t b_len);

// Root tables shared by every caller (grow-only).
bool ensure_tables(uint64_t need, hipStream_t s, const double** fwd, const double** inv, uint64_t* stride);

// ---------------------------------------------------------------------------
// Device memory.

struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    bool ensure(size_t n) {
        if (n <= bytes) return true;
        if (ptr) hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
        size_t want = std::max<size_t>(n, 4096);
        if (hipMalloc(&ptr, want) != hipSuccess) {
            ptr = nullptr;
            return false;
        }
        bytes = want;
        return true;
    }
    void release() {
        if (ptr) hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(ptr);
    }
};

// Pinned (page-locked) host memory, grow-only. By default it is
// non-coherent (HIP_HOST_COHERENT unset): fine for memory that only the DMA
// engines and the host touch, since the GPU may cache it.
struct HostBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    bool ensure(size_t n, unsigned flags = hipHostMallocDefault) {
        if (n <= bytes) return true;
        release();
        size_t want = std::max<size_t>(n, 4096);
        if (hipHostMalloc(&ptr, want, flags) != hipSuccess) {
            ptr = nullptr;
            return false;
        }
        bytes = want;
        return true;
    }
    void release() {
        if (ptr) hipHostFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(ptr);
    }
};

// Pinned host memory mapped into the device's address space: kernels read
// and write it directly over the host link. It is allocated coherent
// (fine-grained): the host rewrites it between calls, so no GPU cache may
// keep a line of it from an earlier kernel (with the default, non-coherent
// memory a later kernel could read stale operand lines and a served request
// came out wrong).
struct MappedBuf {
    HostBuf host;
    void* dev = nullptr;
    bool ensure(size_t n) {
        if (n <= host.bytes && dev) return true;
        dev = nullptr;
        if (!host.ensure(n, hipHostMallocCoherent)) return false;
        if (hipHostGetDevicePointer(&dev, host.ptr, 0) != hipSuccess) {
            dev = nullptr;
            host.release();
            return false;
        }
        return true;
    }
    void release() {
        host.release();
        dev = nullptr;
    }
};

// True when p lies in page-locked host memory (hipHostMalloc), which the
// DMA engines read and write directly.
bool host_pinned(const void* p);

// Copies between pageable host memory and the device through a ring of
// pinned chunk buffers: the host copies one chunk while the DMA engine moves
// another, so the two overlap. Device-to-host copies drain the buffers on
// one thread each (the caller and up to three helpers): the host's copy out
// of DMA-written memory into pageable pages runs at 15-30 GB/s per thread,
// below the link's ~55 GB/s. The runtime's own pageable copies stage through
// buffers too, but at ~15 GB/s device to host (16 MiB, MI355X host) and with
// a first call that maps its staging memory.
struct HostStage {
    static constexpr int kBuffers = 4;
    static constexpr size_t kChunk = (size_t)4 << 20;
    HostBuf buf[kBuffers];
    hipEvent_t ev[kBuffers] = {};
    int next = 0;  // buffer of the next host-to-device chunk
    bool init();
    // Host to device, queued on s; returns once every chunk has been copied
    // out of `host` (the last DMAs may still run: later calls wait for them).
    bool to_device(hipStream_t s, void* dev, const void* host, size_t bytes);
    // Device to host after the work queued on s; returns when `host` holds
    // every byte.
    bool to_host(hipStream_t s, void* host, const void* dev, size_t bytes);
    void release();
    ~HostStage() { release(); }
};

// Host copies of an executor's calls: straight DMA for pinned host memory,
// the staging ring otherwise (and straight copies when the ring cannot be
// allocated).
bool copy_to_device(HostStage& st, hipStream_t s, void* dev, const void* host, size_t bytes);
bool copy_to_host(HostStage& st, hipStream_t s, void* host, const void* dev, size_t bytes);

struct Executor {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    DevBuf spec_a, spec_b, in_a, in_b, out, carry;
    HostStage stage;
    MappedBuf mapped;  // operands and result of small host-buffer calls
    std::mutex lock;

    bool init() {
        if (stream) return true;
        if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) {
            stream = nullptr;
            return false;
        }
        own_stream = true;
        return true;
    }
    ~Executor() {
        spec_a.release();
        spec_b.release();
        in_a.release();
        in_b.release();
        out.release();
        carry.release();
        mapped.release();
        if (own_stream && stream) hipStreamDestroy(stream);
    }
};

Executor* plan_executor();

// Working memory of one operation (see op_spectra_bytes, op_carry_bytes).
struct OpBuffers {
    double* spec_a;
    double* spec_b;  // unused for a square
    uint32_t* carry;
};
// Spectra of operand a (every block of a split's longer operand) and of
// operand b (the shorter operand of a split).
uint64_t op_spectra_bytes(const Geometry& g);
uint64_t op_spectra_bytes_b(const Geometry& g);
// Carry blocks of an output (kSchoolBlockLimbs limbs for a schoolbook
// product, kCrtBlockLimbs otherwise) and their bookkeeping bytes.
uint64_t carry_blocks(uint64_t out_len, bool school = false);
uint64_t op_carry_bytes(uint64_t out_len, bool school = false);

// Whole operations on device operands, queued on stream s.
bool product_on(hipStream_t s, const Geometry& g, const uint64_t* d_a, uint64_t la, const uint64_t* d_b,
                uint64_t lb, bool square, const OpBuffers& buf, uint64_t* d_out);
bool convolution_on(hipStream_t s, const Geometry& g, const ntt_conv_plan* p, const uint64_t* d_a,
                    const uint64_t* d_b, bool square, const OpBuffers& buf, uint64_t* d_out);

// ---------------------------------------------------------------------------
// Transform splits and batched products.

struct Split {
    int n;
    int logs[3];  // strided passes, then the row pass (last)
    // Kernel forms the autotuner may choose beside the levels: the
    // alternative strided tiles (alt_strided_lines), radix-16 inverse rows
    // in the product row kernel, packing per (tile, prime) instead of
    // packing shared across the primes of a tile, and radix-8 row rounds
    // with twice the threads (taking precedence over inv16); strided passes
    // of 2^5, 2^6 or 2^9 points in radix-8 rounds with twice the threads
    // (as many rounds as radix 16 there); product rows with one thread per
    // group of the product, taking both operands' groups (k_fused FORM 3).
    // Strided passes of 2^5 or 2^6 points in wave-owned lines (no workgroup
    // barrier, k_strided_waves).
    uint8_t alt_tiles, inv16, per_prime_pack, rows8, strided8, paired, waves;
    int fused_form() const { return rows8 ? 2 : (paired ? 3 : (inv16 ? 1 : 0)); }
};

// Which tuned split a transform takes: a single operation's product (two
// operands), its square (one operand), or a batch bucket.
enum SplitUse { kUseProduct = 0, kUseSquare = 1, kUseBatch = 2 };
constexpr uint32_t kFormAltTiles = NTT_FORM_ALT_TILES, kFormInv16 = NTT_FORM_INV16,
                   kFormPerPrimePack = NTT_FORM_PER_PRIME_PACK, kFormRows8 = NTT_FORM_ROWS8,
                   kFormStrided8 = NTT_FORM_STRIDED8, kFormPaired = NTT_FORM_PAIRED,
                   kFormWaves = NTT_FORM_WAVES,
                   kFormTuned = NTT_FORM_TUNED;
Split split_of(int K, int use);

struct BatchLaunch {
    enum Kind { kFwdStrided, kPackPrimes, kFused, kInvStrided, kCrt, kSchool } kind;
    int logl;
    bool flag;  // pack (forward strided), canonical output (inverse strided), square (fused)
    bool pack;  // fused single pass: pack the operands
    bool alt;   // strided passes: alternative tiles
    int form;   // fused passes: Split::fused_form
    int units;
    const char* name;
    PassArgs pass;
    CrtArgs crt;
};

struct BatchBucket {
    int K;
    uint32_t fwd, n_fwd;  // forward items: every first operand, then every second one
    uint32_t pk, n_pk;    // the same with prime 0 only: shared packing items
    bool shared_ok;       // every operand's coefficients fit the shared packing pass
    uint32_t ns, n_ns;    // product items of products with two operands
    uint32_t sq, n_sq;    // product items of squares (follow ns: ns..sq is every product)
};

struct Batch {
    Executor ex;
    size_t count = 0;
    std::vector<uint64_t> la, lb;
    std::vector<uint8_t> square;
    std::vector<uint8_t> swapped;  // split product whose operand b is the longer one
    std::vector<Geometry> geo;
    std::vector<BatchProduct> host;
    std::vector<BatchBucket> buckets;
    // Reconstruction items per prime count; index 0 holds the blocks of the
    // schoolbook products.
    uint32_t crt_begin[kMaxPrimes + 1] = {0}, crt_count[kMaxPrimes + 1] = {0};
    // Products of each prime count whose carries take the scan (more than
    // kCarryDirectBlocks output blocks): one item per chunk of carry blocks.
    uint32_t scan_begin[kMaxPrimes + 1] = {0}, scan_count[kMaxPrimes + 1] = {0};
    DevBuf products, items, arena, carry;
    std::vector<BatchLaunch> program;
    // program[seg[i], seg[i+1]) for i < seg.size() - 1 are independent
    // segments (the launches of one bucket, or the schoolbook launch); the
    // reconstruction launches from seg.back() on depend on all of them.
    std::vector<size_t> seg;
    std::vector<float> seg_ms;  // measured duration of each independent segment (empty: not measured)
    std::vector<hipStream_t> side;  // branches of the captured graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool bound = false;
    bool tuned = false;
    const double *fwd = nullptr, *inv = nullptr;
    uint64_t tstride = 0;

    ~Batch() {
        drop_graph();
        for (hipStream_t x : side) hipStreamDestroy(x);
        products.release();
        items.release();
        arena.release();
        carry.release();
    }
    void drop_graph() {
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
};

bool batch_plan(Batch& B, size_t count, const size_t* la, const size_t* lb, const uint8_t* sq);
bool batch_bind(Batch& B, const uint64_t* const* d_a, const uint64_t* const* d_b, uint64_t* const* d_out);
// Queues the batch on `stream` (the batch's own stream when null), as one
// graph replay when use_graph is set, else as plain launches.
bool batch_run(Batch& B, hipStream_t stream, bool use_graph);

}  // namespace nttamd
