// api.cpp -- the C ABI of libmwgemm_hip.so: the reference's Fortran-ABI GEMM entry points
// (mgemm_/mgemm_64_/wgemm_/wgemm_64_/mgemm_serial/wgemm_serial), host<->device staging,
// the numeric-path switch, error reporting and kernel timers.
//
// Semantics follow the reference driver (src/epblas-parallel/multifloats/
// mgemm_parallel.cpp:44-187, wgemm_parallel.cpp:44-194):
//   trans letters upper-cased, real 'C' folded to 'T'       (common/blas_char.h:33-39)
//   m <= 0 or n <= 0: return, C untouched                  (mgemm_parallel.cpp:66)
//   beta pre-pass: beta == 0 -> exact zeros (NaN cleared), beta == 1 -> skip,
//                  else C*beta                              (:68-77)
//   alpha == 0 or k == 0: only the pre-pass                 (:78)
//   trans / uplo letters outside N/T/C and U/L are read as the reference's packers read
//   them, not rejected (see the ref_* letter functions below)
// Deviations, all toward failing loudly instead of silently: leading dimensions shorter
// than the rows they hold are rejected (the reference reads overlapping columns), and a
// device failure is reported and aborts (or, with MWGEMM_ON_ERROR=return, leaves C
// unmodified and is readable through mwgemm_last_error) where the reference's serial path
// would silently leave C = beta*C (mgemm_serial.cpp:353-365).  An exact-product call whose
// workspace cannot be allocated runs the plain path instead, as the reference's parallel
// driver recomputes serially when its pack buffers cannot be had (mgemm_parallel.cpp:
// 180-186).  The device-pointer entry points (no counterpart in the reference) reject
// letters outside N/T/C and U/L.
#include <hip/hip_runtime.h>
#include <sched.h>

#include <algorithm>
#include <atomic>
#include <cctype>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mwgemm_hip.h"
#include "mw_internal.hpp"
#include "host_copy.hpp"
#include "ozaki.hpp"
#include "timer.hpp"

namespace mw {

// ----------------------------------------------------------------------------------
// Errors.
// ----------------------------------------------------------------------------------
namespace {
std::mutex g_err_mu;
int g_err_code = 0;
std::string g_err_msg;
std::atomic<int> g_abort{-1};   // -1: read MWGEMM_ON_ERROR on first use

bool abort_on_error()
{
    int a = g_abort.load();
    if (a < 0) {
        const char *e = getenv("MWGEMM_ON_ERROR");
        a = (e && strcmp(e, "return") == 0) ? 0 : 1;
        g_abort.store(a);
    }
    return a != 0;
}
}  // namespace

int record_error(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    {
        std::lock_guard<std::mutex> lk(g_err_mu);
        g_err_code = code ? code : -1;
        g_err_msg = buf;
    }
    fprintf(stderr, "libmwgemm_hip: error %d: %s\n", code, buf);
    fflush(stderr);
    return code ? code : -1;
}

// A failed HIP call also leaves its error as the thread's last error: clear it, so that the
// launch check of a later call does not report it again (sticky errors persist anyway).
#define MW_TRY(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) {                                                        \
            (void)hipGetLastError();                                                   \
            return record_error((int)e_, "%s failed: %s", #expr, hipGetErrorString(e_)); \
        }                                                                              \
    } while (0)

// ----------------------------------------------------------------------------------
// The library's own device buffers.
// ----------------------------------------------------------------------------------
namespace {
std::atomic<int64_t> g_alloc_limit{0};   // bytes; 0: no limit (test hook)
}

bool device_alloc_allowed(size_t bytes)
{
    const int64_t lim = g_alloc_limit.load();
    return lim <= 0 || bytes <= (size_t)lim;
}

hipError_t device_alloc(void **p, size_t bytes)
{
    *p = nullptr;
    if (!device_alloc_allowed(bytes)) return hipErrorOutOfMemory;
    return hipMalloc(p, bytes);
}

// ----------------------------------------------------------------------------------
// Per-device state: one stream and a grow-only staging/workspace arena.
// ----------------------------------------------------------------------------------
namespace {
struct DevState {
    hipStream_t stream = nullptr;
    void *stage = nullptr;   // host-pointer staging (A | B | C)
    size_t stage_bytes = 0;
    char *pin[2] = {nullptr, nullptr};   // pinned host chunks for host <-> device transfers
    size_t pin_bytes = 0;
    hipEvent_t pin_ev[2] = {nullptr, nullptr};   // last transfer that used each chunk
    bool pin_used[2] = {false, false};
    int pin_slot = 0;                            // chunk the next region is taken from
    size_t pin_off = 0;                          // first free byte in that chunk
    hipEvent_t dl_ev[2] = {nullptr, nullptr};    // completion of download chunks, in turn
    // Column-panel pipeline of large host-pointer GEMMs: its upload and download streams,
    // per-panel events, and pinned chunks of its own for the downloads (the upload ring's
    // per-chunk events name one stream).
    hipStream_t up = nullptr, dn = nullptr;
    hipEvent_t up_ev[64] = {}, comp_ev[64] = {};
    char *dpin[2] = {nullptr, nullptr};
    hipEvent_t dpin_ev[2] = {nullptr, nullptr};
};
std::mutex g_state_mu;
std::deque<DevState> g_dev;   // a deque keeps each device's state at a fixed address
std::mutex g_call_mu;        // host-pointer entry points are serialised
std::atomic<int> g_path{-1};
std::atomic<int> g_order{-1};

int current_device(int *dev)
{
    MW_TRY(hipGetDevice(dev));
    return 0;
}

DevState *dev_state(int dev)
{
    std::lock_guard<std::mutex> lk(g_state_mu);
    if ((int)g_dev.size() <= dev) g_dev.resize(dev + 1);
    return &g_dev[dev];
}

// The host-pointer entry points' stream.  It is a blocking stream: its work is ordered
// after what the caller queued on HIP's default stream (device-pointer operands of a
// drop-in call produced there are complete before they are read).
int ensure_stream(DevState *st)
{
    if (!st->stream) MW_TRY(hipStreamCreateWithFlags(&st->stream, hipStreamDefault));
    return 0;
}

// Device staging for host operands.  Grows by at least half each time, so that a caller
// whose shapes grow call by call (a blocked solver's trailing updates) reallocates rarely.
int ensure_stage(DevState *st, size_t bytes)
{
    if (st->stage_bytes >= bytes && device_alloc_allowed(bytes)) return 0;
    const size_t need = bytes;
    bytes = std::max(bytes, st->stage_bytes + st->stage_bytes / 2);
    bytes = (bytes + (size_t(2) << 20) - 1) & ~((size_t(2) << 20) - 1);
    if (!device_alloc_allowed(bytes)) bytes = need;   // the growth margin is optional
    if (st->stage) {
        MW_TRY(hipStreamSynchronize(st->stream));
        MW_TRY(hipFree(st->stage));
        st->stage = nullptr;
        st->stage_bytes = 0;
    }
    if (device_alloc(&st->stage, bytes) != hipSuccess) {
        (void)hipGetLastError();
        st->stage = nullptr;
        return record_error((int)hipErrorOutOfMemory,
                            "host staging: cannot allocate %zu bytes of device memory", bytes);
    }
    st->stage_bytes = bytes;
    return 0;
}

// Pinned staging: two chunks used as a ring of regions.  Pageable operands are copied (in
// parallel, on the host) into a region while the DMAs of earlier regions are in flight;
// small operands share a chunk, so a small call waits once, for its result.
constexpr size_t PIN_CHUNK = size_t(32) << 20;

size_t pin_align(size_t x) { return (x + 255) & ~size_t(255); }

int wait_slot(DevState *st, int slot)
{
    if (st->pin_used[slot]) MW_TRY(hipEventSynchronize(st->pin_ev[slot]));
    st->pin_used[slot] = false;
    return 0;
}

// Chunks of at least bytes each (PIN_CHUNK unless one region needs more).
int ensure_pinned(DevState *st, size_t bytes)
{
    bytes = pin_align(bytes);
    for (int i = 0; i < 2; ++i) {
        if (!st->pin_ev[i]) MW_TRY(hipEventCreateWithFlags(&st->pin_ev[i], hipEventDisableTiming));
        if (!st->dl_ev[i]) MW_TRY(hipEventCreateWithFlags(&st->dl_ev[i], hipEventDisableTiming));
    }
    if (st->pin_bytes >= bytes) return 0;
    bytes = std::max(bytes, PIN_CHUNK);
    for (int i = 0; i < 2; ++i) {
        if (int rc = wait_slot(st, i)) return rc;
        if (st->pin[i]) {
            MW_TRY(hipHostFree(st->pin[i]));
            st->pin[i] = nullptr;
        }
    }
    st->pin_bytes = 0;
    for (int i = 0; i < 2; ++i)
        MW_TRY(hipHostMalloc(reinterpret_cast<void **>(&st->pin[i]), bytes, hipHostMallocDefault));
    st->pin_bytes = bytes;
    st->pin_slot = 0;
    st->pin_off = 0;
    return 0;
}

// A region of sz bytes (sz <= pin_bytes): the rest of the current chunk if it has room,
// else the other chunk once the transfers that last used it have finished.  The caller
// records pin_ev[pin_slot] after the transfer it issues from the region.
int pin_take(DevState *st, size_t sz, char **out)
{
    sz = pin_align(sz);
    if (st->pin_off + sz > st->pin_bytes) {
        st->pin_slot ^= 1;
        if (int rc = wait_slot(st, st->pin_slot)) return rc;
        st->pin_off = 0;
    }
    *out = st->pin[st->pin_slot] + st->pin_off;
    st->pin_off += sz;
    return 0;
}

int pin_mark(DevState *st, hipStream_t s)
{
    MW_TRY(hipEventRecord(st->pin_ev[st->pin_slot], s));
    st->pin_used[st->pin_slot] = true;
    return 0;
}

// Host (pageable) -> device: cols columns of seg bytes at host pitch hp, packed on the
// device (pitch seg), one contiguous DMA per chunk.
int upload_pinned(DevState *st, hipStream_t s, char *dst, const char *src, size_t hp, size_t seg,
                  int64_t cols)
{
    if (cols <= 0 || seg == 0) return 0;
    const int64_t per = (int64_t)std::max<size_t>(1, PIN_CHUNK / seg);
    if (int rc = ensure_pinned(st, (size_t)std::min<int64_t>(per, cols) * seg)) return rc;
    for (int64_t c0 = 0; c0 < cols; c0 += per) {
        const int64_t nc = std::min(per, cols - c0);
        char *reg = nullptr;
        if (int rc = pin_take(st, (size_t)nc * seg, &reg)) return rc;
        host_copy_columns(reg, seg, src + c0 * hp, hp, seg, nc);
        MW_TRY(hipMemcpyAsync(dst + c0 * seg, reg, (size_t)nc * seg, hipMemcpyHostToDevice, s));
        if (int rc = pin_mark(st, s)) return rc;
    }
    return 0;
}

// One host operand of a staged call: cols columns of seg bytes at host pitch hp, to be
// packed at byte offset doff of the device staging buffer.
struct HostOperand {
    const char *host;
    size_t hp, seg;
    int64_t cols;
    size_t doff;
};

// Uploads the operands.  When their packed span fits one pinned chunk they are gathered
// into one region laid out like the device buffer and travel in a single DMA (a small call
// then costs one transfer each way); otherwise each goes through the chunked pipeline.
int upload_operands(DevState *st, hipStream_t s, char *dbase, const HostOperand *ops, int nops)
{
    size_t lo = SIZE_MAX, hi = 0;
    for (int i = 0; i < nops; ++i) {
        lo = std::min(lo, ops[i].doff);
        hi = std::max(hi, ops[i].doff + ops[i].seg * (size_t)ops[i].cols);
    }
    if (hi <= lo) return 0;
    if (hi - lo <= PIN_CHUNK) {
        if (int rc = ensure_pinned(st, hi - lo)) return rc;
        char *reg = nullptr;
        if (int rc = pin_take(st, hi - lo, &reg)) return rc;
        for (int i = 0; i < nops; ++i)
            host_copy_columns(reg + (ops[i].doff - lo), ops[i].seg, ops[i].host, ops[i].hp,
                              ops[i].seg, ops[i].cols);
        MW_TRY(hipMemcpyAsync(dbase + lo, reg, hi - lo, hipMemcpyHostToDevice, s));
        return pin_mark(st, s);
    }
    for (int i = 0; i < nops; ++i)
        if (int rc = upload_pinned(st, s, dbase + ops[i].doff, ops[i].host, ops[i].hp,
                                   ops[i].seg, ops[i].cols)) return rc;
    return 0;
}

// Packed device columns (pitch seg) -> host (pageable, pitch hp): the DMA of chunk i+1 runs
// while chunk i is copied out.  Chunk i+1's region never overlaps chunk i's: it follows it
// in the same pinned chunk, or lies in the other one, whose earlier regions were all copied
// out before.  Returns once the last chunk is on the host; as the final DMA is the last
// work on the stream, the stream is then idle.
int download_pinned(DevState *st, hipStream_t s, char *dst, size_t hp, const char *src,
                    size_t seg, int64_t cols)
{
    if (cols <= 0 || seg == 0) return 0;
    const int64_t per = (int64_t)std::max<size_t>(1, PIN_CHUNK / seg);
    if (int rc = ensure_pinned(st, (size_t)std::min<int64_t>(per, cols) * seg)) return rc;
    const int64_t nchunk = (cols + per - 1) / per;
    char *reg[2] = {nullptr, nullptr};
    auto issue = [&](int64_t i) -> int {
        const int64_t c0 = i * per, nc = std::min(per, cols - c0);
        if (int rc = pin_take(st, (size_t)nc * seg, &reg[i & 1])) return rc;
        MW_TRY(hipMemcpyAsync(reg[i & 1], src + c0 * seg, (size_t)nc * seg,
                              hipMemcpyDeviceToHost, s));
        MW_TRY(hipEventRecord(st->dl_ev[i & 1], s));
        return pin_mark(st, s);
    };
    if (int rc = issue(0)) return rc;
    for (int64_t i = 0; i < nchunk; ++i) {
        if (i + 1 < nchunk)
            if (int rc = issue(i + 1)) return rc;
        MW_TRY(hipEventSynchronize(st->dl_ev[i & 1]));
        const int64_t c0 = i * per, nc = std::min(per, cols - c0);
        host_copy_columns(dst + c0 * hp, hp, reg[i & 1], seg, seg, nc);
    }
    return 0;
}

int default_path()
{
    int p = g_path.load();
    if (p < 0) {
        const char *e = getenv("MWGEMM_PATH");
        p = MWGEMM_PATH_PLAIN;
        if (e && strcmp(e, "ozaki") == 0) p = MWGEMM_PATH_OZAKI;
        if (e && strcmp(e, "auto") == 0) p = MWGEMM_PATH_AUTO;
        g_path.store(p);
    }
    return p;
}

}  // namespace

// Operation order of the plain path, so that results match the reference build this
// library stands in for: MWGEMM_PLAIN_ORDER ("simd" | "scalar"), else what the reference
// picks on this host.
int plain_order()
{
    int o = g_order.load();
    if (o < 0) {
        const char *e = getenv("MWGEMM_PLAIN_ORDER");
        if (e && strcmp(e, "scalar") == 0) o = ORDER_SCALAR;
        else if (e && strcmp(e, "simd") == 0) o = ORDER_SIMD;
        else o = host_has_avx2_fma() ? ORDER_SIMD : ORDER_SCALAR;
        g_order.store(o);
    }
    return o;
}

namespace {
// Letters of the library's own device entry points: OP_N/OP_T/OP_C, or -1 for any other.
int op_real(char c)
{
    c = (char)toupper((unsigned char)c);
    if (c == 'N') return OP_N;
    if (c == 'T' || c == 'C') return OP_T;
    return -1;
}
int op_cplx(char c)
{
    c = (char)toupper((unsigned char)c);
    if (c == 'N') return OP_N;
    if (c == 'T') return OP_T;
    if (c == 'C') return OP_C;
    return -1;
}

// Letters of the drop-in entry points, read exactly as the reference reads them (it never
// rejects a letter; each packer tests for the letters it knows and treats the rest alike).
//   real GEMM, gemmtr, gemv: anything but 'N' is transposed (blas_char.h:36 folds 'C';
//     the packers test == 'N', mgemm_serial.cpp:52, 71; mgemv.cpp:248-250);
//   complex op(A): 'N', 'T', anything else conjugate-transposed (wgemm_pack_A's final
//     branch, wgemm_serial.cpp:56-73);
//   complex op(B): 'N', 'T', 'C'; any other letter is transposed without conjugation by the
//     AVX2 packer (it conjugates only for 'C', wgemm_serial.cpp:145, 174-185) and
//     conjugate-transposed by the scalar packer (:81-99), so it follows the plain order;
//   complex gemv: 'N', 'C', anything else transposed (wgemv.cpp:308-319);
//   gemmtr uplo: 'U' is the upper triangle, anything else the lower (mgemmtr_serial.cpp:178).
int ref_op_real(char c) { return toupper((unsigned char)c) == 'N' ? OP_N : OP_T; }
int ref_op_cplx_a(char c)
{
    c = (char)toupper((unsigned char)c);
    return c == 'N' ? OP_N : c == 'T' ? OP_T : OP_C;
}
int ref_op_cplx_b(char c, int order)
{
    c = (char)toupper((unsigned char)c);
    if (c == 'N') return OP_N;
    if (c == 'T') return OP_T;
    if (c == 'C') return OP_C;
    return order == ORDER_SCALAR ? OP_C : OP_T;
}
int ref_op_gemv_cplx(char c)
{
    c = (char)toupper((unsigned char)c);
    return c == 'N' ? OP_N : c == 'C' ? OP_C : OP_T;
}
int ref_upper(char c) { return toupper((unsigned char)c) == 'U' ? 1 : 0; }
// The complex gemmtr's diagonal triangles read a letter outside N/T/C as 'T' (transposed
// unless 'N', conjugated only for 'C', wgemmtr_serial.cpp:54-57), while its rectangle, which
// goes through the GEMM's packers, may read it as 'C': bit 0 / bit 1 mark op(A) / op(B)
// taken as 'T' on the diagonal where the rectangle conjugates.
int ref_gemmtr_diag_alt(char ta, char tb, int opa, int opb)
{
    ta = (char)toupper((unsigned char)ta);
    tb = (char)toupper((unsigned char)tb);
    return (opa == OP_C && ta != 'C' ? 1 : 0) | (opb == OP_C && tb != 'C' ? 2 : 0);
}

bool dd_is0(const ddv &x) { return x.hi == 0.0 && x.lo == 0.0; }
bool dd_is1(const ddv &x) { return x.hi == 1.0 && x.lo == 0.0; }
bool zdd_is0(const zddv &x) { return dd_is0(x.re) && dd_is0(x.im); }
bool zdd_is1(const zddv &x) { return dd_is1(x.re) && dd_is0(x.im); }

bool is_device_ptr(const void *p)
{
    if (!p) return false;
    hipPointerAttribute_t attr;
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged ||
           attr.type == hipMemoryTypeUnified;
}

// AUTO takes the exact-product path once the product is large enough to amortise the
// residue split and its depth is one the plans cover.  An explicit OZAKI request beyond
// that depth fails loudly (ozaki_gemm_* report it) instead of silently running PLAIN.
int choose_path(int path, int64_t m, int64_t n, int64_t k, bool cplx)
{
    const bool ok = cplx ? ozaki_supported_cplx(k) : ozaki_supported(k);
    if (path == MWGEMM_PATH_AUTO)
        return ok && (double)m * (double)n * (double)k * (cplx ? 4.0 : 1.0) >= 1.0e9
                   ? MWGEMM_PATH_OZAKI : MWGEMM_PATH_PLAIN;
    return path;
}
}  // namespace

// ----------------------------------------------------------------------------------
// Device-resident cores.
// ----------------------------------------------------------------------------------
// noalloc: the no-allocation entry point (the exact product in the static scratch memory;
// AUTO takes it only where it fits, an explicit OZAKI request that does not fit fails).
int gemm_real_device(GemmArgsR g, int path, hipStream_t s, uint64_t a_tag, bool noalloc)
{
    g.order = plain_order();
    if (g.m <= 0 || g.n <= 0) return 0;
    if (g.opa < 0 || g.opb < 0) return record_error(-2, "mgemm: invalid trans letter");
    const int64_t arows = g.opa == OP_N ? g.m : g.k;
    const int64_t brows = g.opb == OP_N ? g.k : g.n;
    const bool need_ab = g.k > 0 && !dd_is0(g.alpha);
    if (g.ldc < g.m || (need_ab && (g.lda < (arows > 1 ? arows : 1) || g.ldb < (brows > 1 ? brows : 1))))
        return record_error(-3, "mgemm: leading dimension too small (lda=%lld ldb=%lld ldc=%lld)",
                            (long long)g.lda, (long long)g.ldb, (long long)g.ldc);
    g.beta_mode = dd_is0(g.beta) ? BETA_ZERO : dd_is1(g.beta) ? BETA_ONE : BETA_SCALE;
    if (!need_ab) {
        MW_TRY(plain_beta_only_real(g, s));
        return 0;
    }
    if (choose_path(path, g.m, g.n, g.k, false) == MWGEMM_PATH_OZAKI) {
        const int rc = noalloc ? ozaki_gemm_real_noalloc(g, s, path == MWGEMM_PATH_AUTO)
                               : ozaki_gemm_real(g, s, a_tag);
        // Without its workspace the exact product runs on the plain path, which needs none
        // (the reference's driver likewise recomputes serially when its pack buffers cannot
        // be had, mgemm_parallel.cpp:180-186).
        if (rc != OZ_NO_WORKSPACE) return rc;
    }
    MW_TRY(plain_gemm_real(g, s));
    return 0;
}

// Column block of the triangular-output GEMM's diagonal triangles: MGEMMTR_NB /
// WGEMMTR_NB as the reference reads them (mgemmtr_serial.cpp:104-112,
// wgemmtr_serial.cpp), default 8.  The block decides which elements take the scalar path,
// so it is part of the bit-level contract.
int64_t gemmtr_nb(bool cplx)
{
    auto from_env = [](const char *name) -> int64_t {
        const char *e = getenv(name);
        const long v = e ? atol(e) : 0;
        return v > 0 ? (int64_t)v : 8;
    };
    static const int64_t nb_real = from_env("MGEMMTR_NB"), nb_cplx = from_env("WGEMMTR_NB");
    return cplx ? nb_cplx : nb_real;
}

// Rectangle tile of the real triangular-output GEMM: MGEMMTR_TILE=64 (64 x 64) or 32
// (32 x 64); unset: chosen per call.  Tiling does not change any result bit.
int gemmtr_tile()
{
    static const int v = [] {
        const char *e = getenv("MGEMMTR_TILE");
        const long t = e ? atol(e) : 0;
        return t == 64 ? 0 : t == 32 ? 1 : -1;
    }();
    return v;
}

// Uplo of the device entry points: 1 upper, 0 lower, -1 for a letter other than U/L.
int uplo_upper(char uplo)
{
    const char u = (char)toupper((unsigned char)uplo);
    return u == 'U' ? 1 : u == 'L' ? 0 : -1;
}

// Triangular-output GEMM on device memory (g.m == g.n); the plain path always (its result
// is bit-for-bit the reference's).  upper: 1 / 0, or -1 for an invalid uplo letter.
int gemmtr_real_device(GemmArgsR g, int upper, hipStream_t s)
{
    g.order = plain_order();
    if (upper < 0) return record_error(-2, "mgemmtr: invalid uplo letter");
    if (g.opa < 0 || g.opb < 0) return record_error(-2, "mgemmtr: invalid trans letter");
    if (g.n <= 0) return 0;
    const int64_t arows = g.opa == OP_N ? g.n : g.k;
    const int64_t brows = g.opb == OP_N ? g.k : g.n;
    const bool need_ab = g.k > 0 && !dd_is0(g.alpha);
    if (g.ldc < g.n || (need_ab && (g.lda < (arows > 1 ? arows : 1) || g.ldb < (brows > 1 ? brows : 1))))
        return record_error(-3, "mgemmtr: leading dimension too small (lda=%lld ldb=%lld ldc=%lld)",
                            (long long)g.lda, (long long)g.ldb, (long long)g.ldc);
    g.beta_mode = dd_is0(g.beta) ? BETA_ZERO : dd_is1(g.beta) ? BETA_ONE : BETA_SCALE;
    g.variant = gemmtr_tile();
    MW_TRY(plain_gemmtr_real(g, upper == 1, gemmtr_nb(false), s));
    return 0;
}

int gemmtr_cplx_device(GemmArgsZ g, int upper, hipStream_t s)
{
    g.order = plain_order();
    if (upper < 0) return record_error(-2, "wgemmtr: invalid uplo letter");
    if (g.opa < 0 || g.opb < 0) return record_error(-2, "wgemmtr: invalid trans letter");
    if (g.n <= 0) return 0;
    const int64_t arows = g.opa == OP_N ? g.n : g.k;
    const int64_t brows = g.opb == OP_N ? g.k : g.n;
    const bool need_ab = g.k > 0 && !zdd_is0(g.alpha);
    if (g.ldc < g.n || (need_ab && (g.lda < (arows > 1 ? arows : 1) || g.ldb < (brows > 1 ? brows : 1))))
        return record_error(-3, "wgemmtr: leading dimension too small (lda=%lld ldb=%lld ldc=%lld)",
                            (long long)g.lda, (long long)g.ldb, (long long)g.ldc);
    g.beta_mode = zdd_is0(g.beta) ? BETA_ZERO : zdd_is1(g.beta) ? BETA_ONE : BETA_SCALE;
    MW_TRY(plain_gemmtr_cplx(g, upper == 1, gemmtr_nb(true), s));
    return 0;
}

int gemm_cplx_device(GemmArgsZ g, int path, hipStream_t s, uint64_t a_tag)
{
    g.order = plain_order();
    if (g.m <= 0 || g.n <= 0) return 0;
    if (g.opa < 0 || g.opb < 0) return record_error(-2, "wgemm: invalid trans letter");
    const int64_t arows = g.opa == OP_N ? g.m : g.k;
    const int64_t brows = g.opb == OP_N ? g.k : g.n;
    const bool need_ab = g.k > 0 && !zdd_is0(g.alpha);
    if (g.ldc < g.m || (need_ab && (g.lda < (arows > 1 ? arows : 1) || g.ldb < (brows > 1 ? brows : 1))))
        return record_error(-3, "wgemm: leading dimension too small (lda=%lld ldb=%lld ldc=%lld)",
                            (long long)g.lda, (long long)g.ldb, (long long)g.ldc);
    g.beta_mode = zdd_is0(g.beta) ? BETA_ZERO : zdd_is1(g.beta) ? BETA_ONE : BETA_SCALE;
    if (!need_ab) {
        MW_TRY(plain_beta_only_cplx(g, s));
        return 0;
    }
    if (choose_path(path, g.m, g.n, g.k, true) == MWGEMM_PATH_OZAKI) {
        const int rc = ozaki_gemm_cplx(g, s, a_tag);
        if (rc != OZ_NO_WORKSPACE) return rc;   // as for the real GEMM
    }
    MW_TRY(plain_gemm_cplx(g, s));
    return 0;
}

// ----------------------------------------------------------------------------------
// Host-pointer staging shared by the real and complex drop-in entry points.
// ----------------------------------------------------------------------------------
namespace {
// Large host-pointer GEMMs on the plain path, in column panels of C.  A goes up first;
// then, for each panel, its columns of op(B) (and of C when beta != 0) go up on the upload
// stream while the previous panel computes, and each finished panel of C comes back on the
// download stream while the next one computes.  A panel is a whole set of columns, and
// every element of C is computed by the same per-element sequence as in one call, so the
// result is bitwise that of one call.  All of A, B and C are host memory here.
constexpr int MAX_PANELS = 64;

int ensure_panel_state(DevState *st)
{
    if (!st->up) MW_TRY(hipStreamCreateWithFlags(&st->up, hipStreamDefault));
    if (!st->dn) MW_TRY(hipStreamCreateWithFlags(&st->dn, hipStreamDefault));
    for (int i = 0; i < MAX_PANELS; ++i) {
        if (!st->up_ev[i]) MW_TRY(hipEventCreateWithFlags(&st->up_ev[i], hipEventDisableTiming));
        if (!st->comp_ev[i]) MW_TRY(hipEventCreateWithFlags(&st->comp_ev[i], hipEventDisableTiming));
    }
    for (int i = 0; i < 2; ++i) {
        if (!st->dpin_ev[i]) MW_TRY(hipEventCreateWithFlags(&st->dpin_ev[i], hipEventDisableTiming));
        if (!st->dpin[i])
            MW_TRY(hipHostMalloc(reinterpret_cast<void **>(&st->dpin[i]), PIN_CHUNK, hipHostMallocDefault));
    }
    return 0;
}

// Packed device columns (pitch seg) -> host (pitch hp) through the download chunks, on
// stream dn: the DMA of chunk i+1 runs while chunk i is copied out.  Returns once the last
// chunk is on the host.
int download_panel(DevState *st, char *dst, size_t hp, const char *src, size_t seg, int64_t cols)
{
    const int64_t per = (int64_t)std::max<size_t>(1, PIN_CHUNK / seg);
    if (seg > PIN_CHUNK) return record_error(-6, "host pipeline: a column of %zu bytes exceeds the pinned chunk", seg);
    const int64_t nchunk = (cols + per - 1) / per;
    auto issue = [&](int64_t i) -> int {
        const int64_t c0 = i * per, nc = std::min(per, cols - c0);
        MW_TRY(hipMemcpyAsync(st->dpin[i & 1], src + c0 * seg, (size_t)nc * seg,
                              hipMemcpyDeviceToHost, st->dn));
        MW_TRY(hipEventRecord(st->dpin_ev[i & 1], st->dn));
        return 0;
    };
    if (int rc = issue(0)) return rc;
    for (int64_t i = 0; i < nchunk; ++i) {
        // chunk i+1 reuses the buffer of chunk i-1, already copied out
        if (i + 1 < nchunk)
            if (int rc = issue(i + 1)) return rc;
        MW_TRY(hipEventSynchronize(st->dpin_ev[i & 1]));
        const int64_t c0 = i * per, nc = std::min(per, cols - c0);
        host_copy_columns(dst + c0 * hp, hp, st->dpin[i & 1], seg, seg, nc);
    }
    return 0;
}

// b_trans: op(B) is B^T or B^H, so a panel of op(B)'s columns is a band of B's rows.
template <typename T, typename Args, typename Run>
int staged_panels(Args g, int64_t a_rows, int64_t a_cols, bool b_trans, bool beta_zero, Run run)
{
    const size_t sz = sizeof(T);
    const int64_t m = g.m, n = g.n, k = g.k;
    const int64_t nb = std::max<int64_t>(256, ((n + 7) / 8 + 63) / 64 * 64);
    const int64_t npan = (n + nb - 1) / nb;
    if (npan > MAX_PANELS) return record_error(-3, "host pipeline: too many panels");
    std::lock_guard<std::mutex> call_lk(g_call_mu);
    int dev = 0;
    if (int rc = current_device(&dev)) return rc;
    DevState *st = dev_state(dev);
    if (int rc = ensure_stream(st)) return rc;
    if (int rc = ensure_panel_state(st)) return rc;
    hipStream_t s = st->stream;
    const size_t offA = 0;
    const size_t offB = offA + pin_align((size_t)(a_rows * a_cols) * sz);
    const size_t offC = offB + pin_align((size_t)(k * n) * sz);
    if (int rc = ensure_stage(st, offC + pin_align((size_t)(m * n) * sz))) return rc;
    char *base = static_cast<char *>(st->stage);
    const char *hA = reinterpret_cast<const char *>(g.A);
    const char *hB = reinterpret_cast<const char *>(g.B);
    char *hC = reinterpret_cast<char *>(g.C);
    const size_t lda = (size_t)g.lda * sz, ldb = (size_t)g.ldb * sz, ldc = (size_t)g.ldc * sz;
    auto fail = [&](int rc) {
        (void)hipStreamSynchronize(st->up);
        (void)hipStreamSynchronize(s);
        (void)hipStreamSynchronize(st->dn);
        return rc;
    };
    // panel p: upload its operands, then queue its computation behind them
    auto issue = [&](int64_t p) -> int {
        const int64_t j0 = p * nb, w = std::min(nb, n - j0);
        char *dB = base + offB + (size_t)(j0 * k) * sz;
        char *dC = base + offC + (size_t)(j0 * m) * sz;
        if (b_trans) {
            if (int rc = upload_pinned(st, st->up, dB, hB + (size_t)j0 * sz, ldb, (size_t)w * sz, k)) return rc;
        } else {
            if (int rc = upload_pinned(st, st->up, dB, hB + (size_t)j0 * ldb, ldb, (size_t)k * sz, w)) return rc;
        }
        if (!beta_zero)
            if (int rc = upload_pinned(st, st->up, dC, hC + (size_t)j0 * ldc, ldc, (size_t)m * sz, w)) return rc;
        MW_TRY(hipEventRecord(st->up_ev[p], st->up));
        MW_TRY(hipStreamWaitEvent(s, st->up_ev[p], 0));
        Args gp = g;
        gp.n = w;
        gp.A = reinterpret_cast<decltype(gp.A)>(base + offA);
        gp.lda = a_rows;
        gp.B = reinterpret_cast<decltype(gp.B)>(dB);
        gp.ldb = b_trans ? w : k;
        gp.C = reinterpret_cast<decltype(gp.C)>(dC);
        gp.ldc = m;
        if (int rc = run(gp, s)) return rc;
        MW_TRY(hipEventRecord(st->comp_ev[p], s));
        return 0;
    };
    if (int rc = upload_pinned(st, st->up, base + offA, hA, lda, (size_t)a_rows * sz, a_cols)) return fail(rc);
    if (int rc = issue(0)) return fail(rc);
    for (int64_t p = 0; p < npan; ++p) {
        if (p + 1 < npan)
            if (int rc = issue(p + 1)) return fail(rc);
        const int64_t j0 = p * nb, w = std::min(nb, n - j0);
        if (hipStreamWaitEvent(st->dn, st->comp_ev[p], 0) != hipSuccess) {
            (void)hipGetLastError();
            return fail(record_error(-1, "host pipeline: stream wait failed"));
        }
        if (int rc = download_panel(st, hC + (size_t)j0 * ldc, ldc,
                                    base + offC + (size_t)(j0 * m) * sz, (size_t)m * sz, w))
            return fail(rc);
    }
    return 0;
}

// A tag unique to one panelled host call (never 0).
uint64_t next_panel_tag()
{
    static std::atomic<uint64_t> next{0};
    return ++next;
}

// Whether a host-pointer GEMM goes through the column-panel pipeline: A, B and C all in
// host memory, a product of at least 2^30 multiply-adds, at least two 256-column panels.
bool panel_worthy(const void *a, const void *b, const void *c, int64_t m, int64_t n, int64_t k,
                  bool need_ab)
{
    if (!need_ab || n < 512 || (double)m * (double)n * (double)k < 1073741824.0) return false;
    if (getenv("MWGEMM_HOST_PANELS") && atoi(getenv("MWGEMM_HOST_PANELS")) == 0) return false;
    return !is_device_ptr(a) && !is_device_ptr(b) && !is_device_ptr(c);
}

template <typename T, typename Args, typename Run>
int staged_call(Args g, int64_t a_rows, int64_t a_cols, int64_t b_rows, int64_t b_cols,
                bool beta_zero, Run run)
{
    if (g.m <= 0 || g.n <= 0) return 0;
    const int64_t a_elems = a_rows > 0 && a_cols > 0 ? a_rows * a_cols : 0;
    const int64_t b_elems = b_rows > 0 && b_cols > 0 ? b_rows * b_cols : 0;
    const int64_t c_elems = g.m * g.n;
    std::lock_guard<std::mutex> call_lk(g_call_mu);
    int dev = 0;
    if (int rc = current_device(&dev)) return rc;
    DevState *st = dev_state(dev);
    if (int rc = ensure_stream(st)) return rc;
    hipStream_t s = st->stream;

    const T *hA = reinterpret_cast<const T *>(g.A);
    const T *hB = reinterpret_cast<const T *>(g.B);
    T *hC = reinterpret_cast<T *>(g.C);
    const int64_t ldc_host = g.ldc;
    const bool dA = a_elems == 0 || is_device_ptr(hA);
    const bool dB = b_elems == 0 || is_device_ptr(hB);
    const bool dC = is_device_ptr(hC);
    size_t need = 0;
    size_t offA = need; if (!dA) need += pin_align(a_elems * sizeof(T));
    size_t offB = need; if (!dB) need += pin_align(b_elems * sizeof(T));
    size_t offC = need; if (!dC) need += pin_align(c_elems * sizeof(T));
    if (need) {
        if (int rc = ensure_stage(st, need)) return rc;
    }
    char *base = static_cast<char *>(st->stage);
    // Only the logical rows x cols of each operand travel.  The device copies are packed
    // (leading dimension = rows; the kernels' per-element arithmetic does not depend on it),
    // so each transfer is contiguous, and C's padding rows on the host are never written.
    HostOperand ops[3];
    int nops = 0;
    if (!dA) {
        ops[nops++] = {reinterpret_cast<const char *>(hA), g.lda * sizeof(T), a_rows * sizeof(T),
                       a_cols, offA};
        g.A = reinterpret_cast<decltype(g.A)>(base + offA);
        g.lda = a_rows;
    }
    if (!dB) {
        ops[nops++] = {reinterpret_cast<const char *>(hB), g.ldb * sizeof(T), b_rows * sizeof(T),
                       b_cols, offB};
        g.B = reinterpret_cast<decltype(g.B)>(base + offB);
        g.ldb = b_rows;
    }
    if (!dC) {
        if (!beta_zero)
            ops[nops++] = {reinterpret_cast<const char *>(hC), ldc_host * sizeof(T),
                           g.m * sizeof(T), g.n, offC};
        g.C = reinterpret_cast<decltype(g.C)>(base + offC);
        g.ldc = g.m;
    }
    if (int rc = upload_operands(st, s, base, ops, nops)) {
        (void)hipStreamSynchronize(s);
        return rc;
    }
    if (int rc = run(g, s)) {
        (void)hipStreamSynchronize(s);
        return rc;
    }
    if (!dC)
        return download_pinned(st, s, reinterpret_cast<char *>(hC), ldc_host * sizeof(T),
                               base + offC, g.m * sizeof(T), g.n);
    MW_TRY(hipStreamSynchronize(s));
    return 0;
}

int mgemm_host(char transa, char transb, int64_t m, int64_t n, int64_t k,
               const mw_float64x2 *alpha, const mw_float64x2 *a, int64_t lda,
               const mw_float64x2 *b, int64_t ldb, const mw_float64x2 *beta, mw_float64x2 *c,
               int64_t ldc)
{
    if (m <= 0 || n <= 0) return 0;
    GemmArgsR g{};
    g.m = m; g.n = n; g.k = k;
    g.opa = ref_op_real(transa); g.opb = ref_op_real(transb);
    g.alpha = ddv{alpha->limbs[0], alpha->limbs[1]};
    g.beta = ddv{beta->limbs[0], beta->limbs[1]};
    g.A = reinterpret_cast<const ddv *>(a); g.lda = lda;
    g.B = reinterpret_cast<const ddv *>(b); g.ldb = ldb;
    g.C = reinterpret_cast<ddv *>(c); g.ldc = ldc;
    g.host_call = true;
    const bool need_ab = k > 0 && !dd_is0(g.alpha);
    const int64_t ar = need_ab ? (g.opa == OP_N ? m : k) : 0, ac = g.opa == OP_N ? k : m;
    const int64_t br = need_ab ? (g.opb == OP_N ? k : n) : 0, bc = g.opb == OP_N ? n : k;
    if (ldc < m || (need_ab && (lda < ar || ldb < br)))
        return record_error(-3, "mgemm: leading dimension too small (lda=%lld ldb=%lld ldc=%lld)",
                            (long long)lda, (long long)ldb, (long long)ldc);
    const int path = default_path();
    const int eff = choose_path(path, m, n, k, false);
    if (m * (size_t)sizeof(ddv) <= PIN_CHUNK && panel_worthy(a, b, c, m, n, k, need_ab)) {
        // The path is fixed for the whole call; the exact-product panels share op(A)'s
        // maxima and residues through a tag unique to this call.
        const uint64_t tag = eff == MWGEMM_PATH_OZAKI ? next_panel_tag() : 0;
        return staged_panels<ddv>(g, ar, ac, g.opb != OP_N, dd_is0(g.beta),
                                  [eff, tag](GemmArgsR gg, hipStream_t s) { return gemm_real_device(gg, eff, s, tag); });
    }
    return staged_call<ddv>(g, ar, ac, br, bc, dd_is0(g.beta),
                            [path](GemmArgsR gg, hipStream_t s) { return gemm_real_device(gg, path, s); });
}

int wgemm_host(char transa, char transb, int64_t m, int64_t n, int64_t k,
               const mw_complex64x2 *alpha, const mw_complex64x2 *a, int64_t lda,
               const mw_complex64x2 *b, int64_t ldb, const mw_complex64x2 *beta,
               mw_complex64x2 *c, int64_t ldc)
{
    if (m <= 0 || n <= 0) return 0;
    GemmArgsZ g{};
    g.m = m; g.n = n; g.k = k;
    g.opa = ref_op_cplx_a(transa); g.opb = ref_op_cplx_b(transb, plain_order());
    g.alpha = zddv{{alpha->re.limbs[0], alpha->re.limbs[1]}, {alpha->im.limbs[0], alpha->im.limbs[1]}};
    g.beta = zddv{{beta->re.limbs[0], beta->re.limbs[1]}, {beta->im.limbs[0], beta->im.limbs[1]}};
    g.A = reinterpret_cast<const zddv *>(a); g.lda = lda;
    g.B = reinterpret_cast<const zddv *>(b); g.ldb = ldb;
    g.C = reinterpret_cast<zddv *>(c); g.ldc = ldc;
    const bool need_ab = k > 0 && !zdd_is0(g.alpha);
    const int64_t ar = need_ab ? (g.opa == OP_N ? m : k) : 0, ac = g.opa == OP_N ? k : m;
    const int64_t br = need_ab ? (g.opb == OP_N ? k : n) : 0, bc = g.opb == OP_N ? n : k;
    if (ldc < m || (need_ab && (lda < ar || ldb < br)))
        return record_error(-3, "wgemm: leading dimension too small (lda=%lld ldb=%lld ldc=%lld)",
                            (long long)lda, (long long)ldb, (long long)ldc);
    const int path = default_path();
    const int eff = choose_path(path, m, n, k, true);
    if (m * (size_t)sizeof(zddv) <= PIN_CHUNK && panel_worthy(a, b, c, m, n, k, need_ab)) {
        const uint64_t tag = eff == MWGEMM_PATH_OZAKI ? next_panel_tag() : 0;
        return staged_panels<zddv>(g, ar, ac, g.opb != OP_N, zdd_is0(g.beta),
                                   [eff, tag](GemmArgsZ gg, hipStream_t s) { return gemm_cplx_device(gg, eff, s, tag); });
    }
    return staged_call<zddv>(g, ar, ac, br, bc, zdd_is0(g.beta),
                             [path](GemmArgsZ gg, hipStream_t s) { return gemm_cplx_device(gg, path, s); });
}

// Arguments of a triangular-output GEMM; ref: read the letters as the reference does (the
// drop-in entry points), else strictly (the device entry points).
GemmArgsR real_args(char transa, char transb, int64_t m, int64_t n, int64_t k,
                    const mw_float64x2 *alpha, const mw_float64x2 *a, int64_t lda,
                    const mw_float64x2 *b, int64_t ldb, const mw_float64x2 *beta,
                    mw_float64x2 *c, int64_t ldc, bool ref)
{
    GemmArgsR g{};
    g.m = m; g.n = n; g.k = k;
    g.opa = ref ? ref_op_real(transa) : op_real(transa);
    g.opb = ref ? ref_op_real(transb) : op_real(transb);
    g.alpha = ddv{alpha->limbs[0], alpha->limbs[1]};
    g.beta = ddv{beta->limbs[0], beta->limbs[1]};
    g.A = reinterpret_cast<const ddv *>(a); g.lda = lda;
    g.B = reinterpret_cast<const ddv *>(b); g.ldb = ldb;
    g.C = reinterpret_cast<ddv *>(c); g.ldc = ldc;
    return g;
}

GemmArgsZ cplx_args(char transa, char transb, int64_t m, int64_t n, int64_t k,
                    const mw_complex64x2 *alpha, const mw_complex64x2 *a, int64_t lda,
                    const mw_complex64x2 *b, int64_t ldb, const mw_complex64x2 *beta,
                    mw_complex64x2 *c, int64_t ldc, bool ref)
{
    GemmArgsZ g{};
    g.m = m; g.n = n; g.k = k;
    if (ref) {
        g.opa = ref_op_cplx_a(transa);
        g.opb = ref_op_cplx_b(transb, plain_order());
        g.diag_alt = ref_gemmtr_diag_alt(transa, transb, g.opa, g.opb);
    } else {
        g.opa = op_cplx(transa);
        g.opb = op_cplx(transb);
    }
    g.alpha = zddv{{alpha->re.limbs[0], alpha->re.limbs[1]}, {alpha->im.limbs[0], alpha->im.limbs[1]}};
    g.beta = zddv{{beta->re.limbs[0], beta->re.limbs[1]}, {beta->im.limbs[0], beta->im.limbs[1]}};
    g.A = reinterpret_cast<const zddv *>(a); g.lda = lda;
    g.B = reinterpret_cast<const zddv *>(b); g.ldb = ldb;
    g.C = reinterpret_cast<zddv *>(c); g.ldc = ldc;
    return g;
}

// Triangular-output GEMM through host (or device) pointers.  C always travels both ways:
// the complementary triangle must come back as it was.
int mgemmtr_host(char uplo, char transa, char transb, int64_t n, int64_t k,
                 const mw_float64x2 *alpha, const mw_float64x2 *a, int64_t lda,
                 const mw_float64x2 *b, int64_t ldb, const mw_float64x2 *beta, mw_float64x2 *c,
                 int64_t ldc)
{
    if (n <= 0) return 0;
    GemmArgsR g = real_args(transa, transb, n, n, k, alpha, a, lda, b, ldb, beta, c, ldc, true);
    const int upper = ref_upper(uplo);
    const bool need_ab = k > 0 && !dd_is0(g.alpha);
    const int64_t ar = need_ab ? (g.opa == OP_N ? n : k) : 0, ac = g.opa == OP_N ? k : n;
    const int64_t br = need_ab ? (g.opb == OP_N ? k : n) : 0, bc = g.opb == OP_N ? n : k;
    if (ldc < n || (need_ab && (lda < ar || ldb < br)))
        return record_error(-3, "mgemmtr: leading dimension too small (lda=%lld ldb=%lld ldc=%lld)",
                            (long long)lda, (long long)ldb, (long long)ldc);
    return staged_call<ddv>(g, ar, ac, br, bc, false,
                            [upper](GemmArgsR gg, hipStream_t s) { return gemmtr_real_device(gg, upper, s); });
}

int wgemmtr_host(char uplo, char transa, char transb, int64_t n, int64_t k,
                 const mw_complex64x2 *alpha, const mw_complex64x2 *a, int64_t lda,
                 const mw_complex64x2 *b, int64_t ldb, const mw_complex64x2 *beta,
                 mw_complex64x2 *c, int64_t ldc)
{
    if (n <= 0) return 0;
    GemmArgsZ g = cplx_args(transa, transb, n, n, k, alpha, a, lda, b, ldb, beta, c, ldc, true);
    const int upper = ref_upper(uplo);
    const bool need_ab = k > 0 && !zdd_is0(g.alpha);
    const int64_t ar = need_ab ? (g.opa == OP_N ? n : k) : 0, ac = g.opa == OP_N ? k : n;
    const int64_t br = need_ab ? (g.opb == OP_N ? k : n) : 0, bc = g.opb == OP_N ? n : k;
    if (ldc < n || (need_ab && (lda < ar || ldb < br)))
        return record_error(-3, "wgemmtr: leading dimension too small (lda=%lld ldb=%lld ldc=%lld)",
                            (long long)lda, (long long)ldb, (long long)ldc);
    return staged_call<zddv>(g, ar, ac, br, bc, false,
                             [upper](GemmArgsZ gg, hipStream_t s) { return gemmtr_cplx_device(gg, upper, s); });
}

// ----------------------------------------------------------------------------------
// Matrix-vector products.
// ----------------------------------------------------------------------------------
// Row bands of the real op-'N' product: the reference splits the rows over its OpenMP team
// when m >= 64 (mgemv.cpp:95-100, at most 256 threads), and the rows after the last whole
// 4-block of each band take the faithful operations, so the team size is part of the bit
// pattern.  MWGEMV_THREADS names it; else the team an OpenMP runtime forms by default
// (omp_get_max_threads): the first entry of OMP_NUM_THREADS, else the number of CPUs in the
// process's affinity mask (what taskset and cgroup cpusets restrict, not the CPUs online),
// either capped by OMP_THREAD_LIMIT.
int64_t omp_default_team()
{
    auto env_int = [](const char *name) -> long {
        const char *e = getenv(name);
        return e ? atol(e) : 0;   // atol reads the first entry of a list such as "8,4"
    };
    long team = env_int("OMP_NUM_THREADS");
    if (team <= 0) {
        cpu_set_t set;
        CPU_ZERO(&set);
        if (sched_getaffinity(0, sizeof set, &set) == 0) team = CPU_COUNT(&set);
        if (team <= 0) {
            const unsigned hw = std::thread::hardware_concurrency();
            team = hw ? (long)hw : 1;
        }
    }
    const long limit = env_int("OMP_THREAD_LIMIT");
    if (limit > 0 && team > limit) team = limit;
    return (int64_t)team;
}

int64_t gemv_bands(int64_t m)
{
    static const int64_t team = [] {
        const char *e = getenv("MWGEMV_THREADS");
        const long v = e ? atol(e) : 0;
        return v > 0 ? (int64_t)v : omp_default_team();
    }();
    if (m < 64) return 1;
    return team < 256 ? team : 256;
}

int gemv_check(const char *name, int op, int64_t m, int64_t n, int64_t lda, int64_t incx,
               int64_t incy)
{
    if (op < 0) return record_error(-2, "%s: invalid trans letter", name);
    if (m < 0 || n < 0) return record_error(-3, "%s: negative dimension", name);
    if (lda < (m > 1 ? m : 1))
        return record_error(-3, "%s: leading dimension too small (lda=%lld)", name, (long long)lda);
    if (incx == 0 || incy == 0) return record_error(-3, "%s: zero vector increment", name);
    return 0;
}

int64_t span_of(int64_t len, int64_t inc) { return len > 0 ? 1 + (len - 1) * (inc < 0 ? -inc : inc) : 0; }

// Host-pointer staging for the matrix-vector products: A's logical rows and the x / y spans
// travel through the pinned chunks; y's span comes back whole (entries between strided
// ones included, unchanged).
template <typename T, typename Run>
int staged_gemv(const T *A, int64_t lda, int64_t m, int64_t n, const T *x, int64_t xspan, T *y,
                int64_t yspan, Run run)
{
    std::lock_guard<std::mutex> call_lk(g_call_mu);
    int dev = 0;
    if (int rc = current_device(&dev)) return rc;
    DevState *st = dev_state(dev);
    if (int rc = ensure_stream(st)) return rc;
    hipStream_t s = st->stream;
    const bool dA = is_device_ptr(A), dx = is_device_ptr(x), dy = is_device_ptr(y);
    size_t need = 0;
    const size_t offA = need; if (!dA) need += pin_align(m * n * sizeof(T));
    const size_t offx = need; if (!dx) need += pin_align(xspan * sizeof(T));
    const size_t offy = need; if (!dy) need += pin_align(yspan * sizeof(T));
    if (need)
        if (int rc = ensure_stage(st, need)) return rc;
    char *base = static_cast<char *>(st->stage);
    const T *gA = A, *gx = x;
    T *gy = y;
    int64_t glda = lda;
    // A travels packed (leading dimension m on the device); the vectors as their whole
    // strided spans, so the kernels see the caller's increments.
    HostOperand ops[3];
    int nops = 0;
    if (!dA) {
        ops[nops++] = {reinterpret_cast<const char *>(A), lda * sizeof(T), m * sizeof(T), n, offA};
        gA = reinterpret_cast<const T *>(base + offA);
        glda = m;
    }
    if (!dx) {
        ops[nops++] = {reinterpret_cast<const char *>(x), xspan * sizeof(T), xspan * sizeof(T), 1,
                       offx};
        gx = reinterpret_cast<const T *>(base + offx);
    }
    if (!dy) {
        ops[nops++] = {reinterpret_cast<const char *>(y), yspan * sizeof(T), yspan * sizeof(T), 1,
                       offy};
        gy = reinterpret_cast<T *>(base + offy);
    }
    if (int rc = upload_operands(st, s, base, ops, nops)) {
        (void)hipStreamSynchronize(s);
        return rc;
    }
    if (int rc = run(gA, glda, gx, gy, s)) {
        (void)hipStreamSynchronize(s);
        return rc;
    }
    if (!dy)
        return download_pinned(st, s, reinterpret_cast<char *>(y), yspan * sizeof(T),
                               base + offy, yspan * sizeof(T), 1);
    MW_TRY(hipStreamSynchronize(s));
    return 0;
}

int mgemv_host(char trans, int64_t m, int64_t n, const mw_float64x2 *alpha,
               const mw_float64x2 *a, int64_t lda, const mw_float64x2 *x, int64_t incx,
               const mw_float64x2 *beta, mw_float64x2 *y, int64_t incy)
{
    const int op = ref_op_real(trans);
    if (int rc = gemv_check("mgemv", op, m, n, lda, incx, incy)) return rc;
    if (m == 0 || n == 0) return 0;
    const bool notrans = op == OP_N;
    const ddv al{alpha->limbs[0], alpha->limbs[1]}, be{beta->limbs[0], beta->limbs[1]};
    const int64_t nb = notrans ? gemv_bands(m) : 1;
    return staged_gemv<ddv>(reinterpret_cast<const ddv *>(a), lda, m, n,
                            reinterpret_cast<const ddv *>(x), span_of(notrans ? n : m, incx),
                            reinterpret_cast<ddv *>(y), span_of(notrans ? m : n, incy),
                            [&](const ddv *gA, int64_t glda, const ddv *gx, ddv *gy, hipStream_t s) {
                                return gemv_real(op, m, n, al, be, gA, glda, gx, incx, gy, incy, nb,
                                                 plain_order(), s);
                            });
}

int wgemv_host(char trans, int64_t m, int64_t n, const mw_complex64x2 *alpha,
               const mw_complex64x2 *a, int64_t lda, const mw_complex64x2 *x, int64_t incx,
               const mw_complex64x2 *beta, mw_complex64x2 *y, int64_t incy)
{
    const int op = ref_op_gemv_cplx(trans);
    if (int rc = gemv_check("wgemv", op, m, n, lda, incx, incy)) return rc;
    if (m == 0 || n == 0) return 0;
    const bool notrans = op == OP_N;
    const zddv al{{alpha->re.limbs[0], alpha->re.limbs[1]}, {alpha->im.limbs[0], alpha->im.limbs[1]}};
    const zddv be{{beta->re.limbs[0], beta->re.limbs[1]}, {beta->im.limbs[0], beta->im.limbs[1]}};
    return staged_gemv<zddv>(reinterpret_cast<const zddv *>(a), lda, m, n,
                             reinterpret_cast<const zddv *>(x), span_of(notrans ? n : m, incx),
                             reinterpret_cast<zddv *>(y), span_of(notrans ? m : n, incy),
                             [&](const zddv *gA, int64_t glda, const zddv *gx, zddv *gy,
                                 hipStream_t s) {
                                 return gemv_cplx(op, m, n, al, be, gA, glda, gx, incx, gy, incy,
                                                  plain_order(), s);
                             });
}

void finish(int rc)
{
    if (rc != 0 && abort_on_error()) {
        fprintf(stderr, "libmwgemm_hip: aborting (set MWGEMM_ON_ERROR=return to get the error back instead)\n");
        abort();
    }
}
}  // namespace

// ----------------------------------------------------------------------------------
// Kernel timers.
// ----------------------------------------------------------------------------------
namespace {
struct TimerSlot { int cls; hipEvent_t a, b; };
std::mutex g_timer_mu;
std::atomic<bool> g_timer_on{false};
std::vector<TimerSlot> g_slots;
double g_timer_ms[KC_COUNT];
int64_t g_timer_cnt[KC_COUNT];
const char *kClassNames[KC_COUNT] = {
    "plain_real", "plain_cplx", "beta_only", "oz_exponent", "oz_split", "oz_gemm", "oz_crt",
    "batch_plain", "batch_oz_split", "batch_oz_gemm", "batch_oz_crt", "gemmtr", "gemv"};

void drain_slots()
{
    for (auto &sl : g_slots) {
        float ms = 0.f;
        if (hipEventSynchronize(sl.b) == hipSuccess && hipEventElapsedTime(&ms, sl.a, sl.b) == hipSuccess) {
            g_timer_ms[sl.cls] += ms;
            g_timer_cnt[sl.cls] += 1;
        }
        (void)hipEventDestroy(sl.a);
        (void)hipEventDestroy(sl.b);
    }
    g_slots.clear();
}
}  // namespace

namespace {
thread_local bool t_timers_paused = false;
}

TimerPause::TimerPause() : prev(t_timers_paused) { t_timers_paused = true; }
TimerPause::~TimerPause() { t_timers_paused = prev; }

int timer_begin(int cls, hipStream_t s)
{
    if (t_timers_paused || !g_timer_on.load(std::memory_order_relaxed)) return -1;
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone) {
        (void)hipGetLastError();
        return -1;    // no event records inside a captured graph
    }
    std::lock_guard<std::mutex> lk(g_timer_mu);
    TimerSlot sl{cls, nullptr, nullptr};
    if (hipEventCreate(&sl.a) != hipSuccess || hipEventCreate(&sl.b) != hipSuccess) return -1;
    if (hipEventRecord(sl.a, s) != hipSuccess) return -1;
    g_slots.push_back(sl);
    return (int)g_slots.size() - 1;
}

void timer_end(int id, hipStream_t s)
{
    if (id < 0) return;
    std::lock_guard<std::mutex> lk(g_timer_mu);
    if (id < (int)g_slots.size()) (void)hipEventRecord(g_slots[id].b, s);
}

}  // namespace mw

using namespace mw;

// ----------------------------------------------------------------------------------
// Exported C ABI.
// ----------------------------------------------------------------------------------
extern "C" {

void mgemm_(const char *transa, const char *transb, const int *m, const int *n, const int *k,
            const mw_float64x2 *alpha, const mw_float64x2 *a, const int *lda,
            const mw_float64x2 *b, const int *ldb, const mw_float64x2 *beta, mw_float64x2 *c,
            const int *ldc)
{
    finish(mgemm_host(*transa, *transb, *m, *n, *k, alpha, a, *lda, b, *ldb, beta, c, *ldc));
}

void mgemm_64_(const char *transa, const char *transb, const int64_t *m, const int64_t *n,
               const int64_t *k, const mw_float64x2 *alpha, const mw_float64x2 *a,
               const int64_t *lda, const mw_float64x2 *b, const int64_t *ldb,
               const mw_float64x2 *beta, mw_float64x2 *c, const int64_t *ldc)
{
    finish(mgemm_host(*transa, *transb, *m, *n, *k, alpha, a, *lda, b, *ldb, beta, c, *ldc));
}

void wgemm_(const char *transa, const char *transb, const int *m, const int *n, const int *k,
            const mw_complex64x2 *alpha, const mw_complex64x2 *a, const int *lda,
            const mw_complex64x2 *b, const int *ldb, const mw_complex64x2 *beta,
            mw_complex64x2 *c, const int *ldc)
{
    finish(wgemm_host(*transa, *transb, *m, *n, *k, alpha, a, *lda, b, *ldb, beta, c, *ldc));
}

void wgemm_64_(const char *transa, const char *transb, const int64_t *m, const int64_t *n,
               const int64_t *k, const mw_complex64x2 *alpha, const mw_complex64x2 *a,
               const int64_t *lda, const mw_complex64x2 *b, const int64_t *ldb,
               const mw_complex64x2 *beta, mw_complex64x2 *c, const int64_t *ldc)
{
    finish(wgemm_host(*transa, *transb, *m, *n, *k, alpha, a, *lda, b, *ldb, beta, c, *ldc));
}

void mgemm_serial(char transa, char transb, ptrdiff_t m, ptrdiff_t n, ptrdiff_t k,
                  const mw_float64x2 *alpha, const mw_float64x2 *a, ptrdiff_t lda,
                  const mw_float64x2 *b, ptrdiff_t ldb, const mw_float64x2 *beta,
                  mw_float64x2 *c, ptrdiff_t ldc)
{
    finish(mgemm_host(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc));
}

void wgemm_serial(char transa, char transb, ptrdiff_t m, ptrdiff_t n, ptrdiff_t k,
                  const mw_complex64x2 *alpha, const mw_complex64x2 *a, ptrdiff_t lda,
                  const mw_complex64x2 *b, ptrdiff_t ldb, const mw_complex64x2 *beta,
                  mw_complex64x2 *c, ptrdiff_t ldc)
{
    finish(wgemm_host(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc));
}

void mgemmtr_(const char *uplo, const char *transa, const char *transb, const int *n,
              const int *k, const mw_float64x2 *alpha, const mw_float64x2 *a, const int *lda,
              const mw_float64x2 *b, const int *ldb, const mw_float64x2 *beta, mw_float64x2 *c,
              const int *ldc)
{
    finish(mgemmtr_host(*uplo, *transa, *transb, *n, *k, alpha, a, *lda, b, *ldb, beta, c, *ldc));
}

void mgemmtr_64_(const char *uplo, const char *transa, const char *transb, const int64_t *n,
                 const int64_t *k, const mw_float64x2 *alpha, const mw_float64x2 *a,
                 const int64_t *lda, const mw_float64x2 *b, const int64_t *ldb,
                 const mw_float64x2 *beta, mw_float64x2 *c, const int64_t *ldc)
{
    finish(mgemmtr_host(*uplo, *transa, *transb, *n, *k, alpha, a, *lda, b, *ldb, beta, c, *ldc));
}

void wgemmtr_(const char *uplo, const char *transa, const char *transb, const int *n,
              const int *k, const mw_complex64x2 *alpha, const mw_complex64x2 *a,
              const int *lda, const mw_complex64x2 *b, const int *ldb,
              const mw_complex64x2 *beta, mw_complex64x2 *c, const int *ldc)
{
    finish(wgemmtr_host(*uplo, *transa, *transb, *n, *k, alpha, a, *lda, b, *ldb, beta, c, *ldc));
}

void wgemmtr_64_(const char *uplo, const char *transa, const char *transb, const int64_t *n,
                 const int64_t *k, const mw_complex64x2 *alpha, const mw_complex64x2 *a,
                 const int64_t *lda, const mw_complex64x2 *b, const int64_t *ldb,
                 const mw_complex64x2 *beta, mw_complex64x2 *c, const int64_t *ldc)
{
    finish(wgemmtr_host(*uplo, *transa, *transb, *n, *k, alpha, a, *lda, b, *ldb, beta, c, *ldc));
}

void mgemmtr_serial(char uplo, char transa, char transb, ptrdiff_t n, ptrdiff_t k,
                    const mw_float64x2 *alpha, const mw_float64x2 *a, ptrdiff_t lda,
                    const mw_float64x2 *b, ptrdiff_t ldb, const mw_float64x2 *beta,
                    mw_float64x2 *c, ptrdiff_t ldc)
{
    finish(mgemmtr_host(uplo, transa, transb, n, k, alpha, a, lda, b, ldb, beta, c, ldc));
}

void wgemmtr_serial(char uplo, char transa, char transb, ptrdiff_t n, ptrdiff_t k,
                    const mw_complex64x2 *alpha, const mw_complex64x2 *a, ptrdiff_t lda,
                    const mw_complex64x2 *b, ptrdiff_t ldb, const mw_complex64x2 *beta,
                    mw_complex64x2 *c, ptrdiff_t ldc)
{
    finish(wgemmtr_host(uplo, transa, transb, n, k, alpha, a, lda, b, ldb, beta, c, ldc));
}

int mwgemm_mgemmtr_device(char uplo, char transa, char transb, int64_t n, int64_t k,
                          const mw_float64x2 *alpha, const mw_float64x2 *dA, int64_t lda,
                          const mw_float64x2 *dB, int64_t ldb, const mw_float64x2 *beta,
                          mw_float64x2 *dC, int64_t ldc, void *stream)
{
    hipStream_t s = static_cast<hipStream_t>(stream);   // NULL: HIP's default stream
    return gemmtr_real_device(real_args(transa, transb, n, n, k, alpha, dA, lda, dB, ldb, beta,
                                        dC, ldc, false), uplo_upper(uplo), s);
}

int mwgemm_wgemmtr_device(char uplo, char transa, char transb, int64_t n, int64_t k,
                          const mw_complex64x2 *alpha, const mw_complex64x2 *dA, int64_t lda,
                          const mw_complex64x2 *dB, int64_t ldb, const mw_complex64x2 *beta,
                          mw_complex64x2 *dC, int64_t ldc, void *stream)
{
    hipStream_t s = static_cast<hipStream_t>(stream);   // NULL: HIP's default stream
    return gemmtr_cplx_device(cplx_args(transa, transb, n, n, k, alpha, dA, lda, dB, ldb, beta,
                                        dC, ldc, false), uplo_upper(uplo), s);
}

void mgemv_(const char *trans, const int *m, const int *n, const mw_float64x2 *alpha,
            const mw_float64x2 *a, const int *lda, const mw_float64x2 *x, const int *incx,
            const mw_float64x2 *beta, mw_float64x2 *y, const int *incy)
{
    finish(mgemv_host(*trans, *m, *n, alpha, a, *lda, x, *incx, beta, y, *incy));
}

void mgemv_64_(const char *trans, const int64_t *m, const int64_t *n, const mw_float64x2 *alpha,
               const mw_float64x2 *a, const int64_t *lda, const mw_float64x2 *x,
               const int64_t *incx, const mw_float64x2 *beta, mw_float64x2 *y,
               const int64_t *incy)
{
    finish(mgemv_host(*trans, *m, *n, alpha, a, *lda, x, *incx, beta, y, *incy));
}

void wgemv_(const char *trans, const int *m, const int *n, const mw_complex64x2 *alpha,
            const mw_complex64x2 *a, const int *lda, const mw_complex64x2 *x, const int *incx,
            const mw_complex64x2 *beta, mw_complex64x2 *y, const int *incy)
{
    finish(wgemv_host(*trans, *m, *n, alpha, a, *lda, x, *incx, beta, y, *incy));
}

void wgemv_64_(const char *trans, const int64_t *m, const int64_t *n,
               const mw_complex64x2 *alpha, const mw_complex64x2 *a, const int64_t *lda,
               const mw_complex64x2 *x, const int64_t *incx, const mw_complex64x2 *beta,
               mw_complex64x2 *y, const int64_t *incy)
{
    finish(wgemv_host(*trans, *m, *n, alpha, a, *lda, x, *incx, beta, y, *incy));
}

int mwgemm_mgemv_device(char trans, int64_t m, int64_t n, const mw_float64x2 *alpha,
                        const mw_float64x2 *dA, int64_t lda, const mw_float64x2 *dx,
                        int64_t incx, const mw_float64x2 *beta, mw_float64x2 *dy, int64_t incy,
                        int64_t nbands, void *stream)
{
    const int op = op_real(trans);
    if (int rc = gemv_check("mgemv", op, m, n, lda, incx, incy)) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);   // NULL: HIP's default stream
    if (nbands <= 0) nbands = op == OP_N ? gemv_bands(m) : 1;
    MW_TRY(gemv_real(op, m, n, ddv{alpha->limbs[0], alpha->limbs[1]},
                     ddv{beta->limbs[0], beta->limbs[1]}, reinterpret_cast<const ddv *>(dA), lda,
                     reinterpret_cast<const ddv *>(dx), incx, reinterpret_cast<ddv *>(dy), incy,
                     nbands, plain_order(), s));
    return 0;
}

int mwgemm_wgemv_device(char trans, int64_t m, int64_t n, const mw_complex64x2 *alpha,
                        const mw_complex64x2 *dA, int64_t lda, const mw_complex64x2 *dx,
                        int64_t incx, const mw_complex64x2 *beta, mw_complex64x2 *dy,
                        int64_t incy, void *stream)
{
    const int op = op_cplx(trans);
    if (int rc = gemv_check("wgemv", op, m, n, lda, incx, incy)) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);   // NULL: HIP's default stream
    const zddv al{{alpha->re.limbs[0], alpha->re.limbs[1]}, {alpha->im.limbs[0], alpha->im.limbs[1]}};
    const zddv be{{beta->re.limbs[0], beta->re.limbs[1]}, {beta->im.limbs[0], beta->im.limbs[1]}};
    MW_TRY(gemv_cplx(op, m, n, al, be, reinterpret_cast<const zddv *>(dA), lda,
                     reinterpret_cast<const zddv *>(dx), incx, reinterpret_cast<zddv *>(dy), incy,
                     plain_order(), s));
    return 0;
}

int mwgemm_set_path(int path)
{
    if (path < MWGEMM_PATH_PLAIN || path > MWGEMM_PATH_AUTO) return -1;
    int prev = default_path();
    g_path.store(path);
    return prev;
}

int mwgemm_get_path(void) { return default_path(); }

int mwgemm_set_plain_order(int order)
{
    if (order != MWGEMM_ORDER_SIMD && order != MWGEMM_ORDER_SCALAR) return -1;
    const int prev = plain_order();
    g_order.store(order);
    return prev;
}

int mwgemm_get_plain_order(void) { return plain_order(); }

int mwgemm_last_error(char *buf, size_t len)
{
    std::lock_guard<std::mutex> lk(g_err_mu);
    if (buf && len) {
        strncpy(buf, g_err_msg.c_str(), len - 1);
        buf[len - 1] = '\0';
    }
    return g_err_code;
}

void mwgemm_clear_error(void)
{
    std::lock_guard<std::mutex> lk(g_err_mu);
    g_err_code = 0;
    g_err_msg.clear();
}

int mwgemm_set_abort(int on)
{
    int prev = abort_on_error() ? 1 : 0;
    g_abort.store(on ? 1 : 0);
    return prev;
}

int mwgemm_mgemm_device(char transa, char transb, int64_t m, int64_t n, int64_t k,
                        const mw_float64x2 *alpha, const mw_float64x2 *dA, int64_t lda,
                        const mw_float64x2 *dB, int64_t ldb, const mw_float64x2 *beta,
                        mw_float64x2 *dC, int64_t ldc, int path, void *stream)
{
    GemmArgsR g{};
    g.m = m; g.n = n; g.k = k;
    g.opa = op_real(transa); g.opb = op_real(transb);
    g.alpha = ddv{alpha->limbs[0], alpha->limbs[1]};
    g.beta = ddv{beta->limbs[0], beta->limbs[1]};
    g.A = reinterpret_cast<const ddv *>(dA); g.lda = lda;
    g.B = reinterpret_cast<const ddv *>(dB); g.ldb = ldb;
    g.C = reinterpret_cast<ddv *>(dC); g.ldc = ldc;
    hipStream_t s = static_cast<hipStream_t>(stream);   // NULL: HIP's default stream
    return gemm_real_device(g, path, s);
}

int mwgemm_mgemm_device_noalloc(char transa, char transb, int64_t m, int64_t n, int64_t k,
                                const mw_float64x2 *alpha, const mw_float64x2 *dA, int64_t lda,
                                const mw_float64x2 *dB, int64_t ldb, const mw_float64x2 *beta,
                                mw_float64x2 *dC, int64_t ldc, int path, void *stream)
{
    TimerPause no_timer_events;
    GemmArgsR g{};
    g.m = m; g.n = n; g.k = k;
    g.opa = op_real(transa); g.opb = op_real(transb);
    g.alpha = ddv{alpha->limbs[0], alpha->limbs[1]};
    g.beta = ddv{beta->limbs[0], beta->limbs[1]};
    g.A = reinterpret_cast<const ddv *>(dA); g.lda = lda;
    g.B = reinterpret_cast<const ddv *>(dB); g.ldb = ldb;
    g.C = reinterpret_cast<ddv *>(dC); g.ldc = ldc;
    if (path < MWGEMM_PATH_PLAIN || path > MWGEMM_PATH_AUTO)
        return record_error(-3, "no-allocation gemm: bad path %d", path);
    hipStream_t s = static_cast<hipStream_t>(stream);   // NULL: HIP's default stream
    return gemm_real_device(g, path, s, 0, true);
}

int mwgemm_wgemm_device(char transa, char transb, int64_t m, int64_t n, int64_t k,
                        const mw_complex64x2 *alpha, const mw_complex64x2 *dA, int64_t lda,
                        const mw_complex64x2 *dB, int64_t ldb, const mw_complex64x2 *beta,
                        mw_complex64x2 *dC, int64_t ldc, int path, void *stream)
{
    GemmArgsZ g{};
    g.m = m; g.n = n; g.k = k;
    g.opa = op_cplx(transa); g.opb = op_cplx(transb);
    g.alpha = zddv{{alpha->re.limbs[0], alpha->re.limbs[1]}, {alpha->im.limbs[0], alpha->im.limbs[1]}};
    g.beta = zddv{{beta->re.limbs[0], beta->re.limbs[1]}, {beta->im.limbs[0], beta->im.limbs[1]}};
    g.A = reinterpret_cast<const zddv *>(dA); g.lda = lda;
    g.B = reinterpret_cast<const zddv *>(dB); g.ldb = ldb;
    g.C = reinterpret_cast<zddv *>(dC); g.ldc = ldc;
    hipStream_t s = static_cast<hipStream_t>(stream);   // NULL: HIP's default stream
    return gemm_cplx_device(g, path, s);
}

int mwgemm_ozaki_config(int64_t k, int cplx, int *nmod, int *bits)
{
    OzPlan p;
    const bool ok = !cplx && ozaki_estimate() != 0
                        ? ozaki_make_fixed_plan(k, ozaki_estimate_bits(), &p)
                        : ozaki_make_plan(cplx ? 2 * k : k, ozaki_target_bits(), ozaki_forced_nmod(), &p);
    if (!ok) return -1;
    if (nmod) *nmod = p.nmod;
    if (bits) *bits = p.bits;
    return 0;
}

int mwgemm_ozaki_set(int bits, int nmod)
{
    if ((bits != 0 && (bits < 8 || bits > 104)) || nmod < 0 || nmod > OZ_MAX_MOD) return -1;
    ozaki_set_params(bits, nmod);
    return 0;
}

int mwgemm_ozaki_set_estimate(int mode)
{
    if (mode < 0 || mode > 2) return -1;
    ozaki_set_estimate(mode);
    return 0;
}

int mwgemm_ozaki_last_moduli(int *nmod)
{
    if (!nmod) return -1;
    return ozaki_last_estimate(nmod);
}

int mwgemm_testing_ozaki_count_offset(int offset)
{
    ozaki_set_count_offset(offset);
    return 0;
}

int mwgemm_ozaki_moduli(int nmod, int *m, int *y)
{
    OzPlan p;
    if (nmod < 1 || nmod > OZ_MAX_MOD || !ozaki_make_plan(1, 8, nmod, &p)) return -1;
    for (int t = 0; t < nmod; ++t) {
        if (m) m[t] = (int)p.mc[t].m;
        if (y) y[t] = (int)p.mc[t].y;
    }
    return 0;
}

int mwgemm_testing_set_alloc_limit(int64_t bytes)
{
    if (bytes < 0) return -1;
    g_alloc_limit.store(bytes);
    return 0;
}

int mwgemm_testing_residue_gemm(const void *dA, const void *dB, void *dU, int64_t mp,
                                int64_t np, int64_t kp, int mod_index, int variant, void *stream)
{
    return ozaki_residue_gemm_test(static_cast<const int8_t *>(dA), static_cast<const int8_t *>(dB),
                                   static_cast<uint8_t *>(dU), mp, np, kp, mod_index, variant,
                                   static_cast<hipStream_t>(stream));
}

int mwgemm_timer_enable(int on)
{
    bool prev = g_timer_on.exchange(on != 0);
    return prev ? 1 : 0;
}

int mwgemm_timer_reset(void)
{
    std::lock_guard<std::mutex> lk(g_timer_mu);
    drain_slots();
    for (int i = 0; i < KC_COUNT; ++i) { g_timer_ms[i] = 0.0; g_timer_cnt[i] = 0; }
    return 0;
}

int mwgemm_timer_read(double *ms, int64_t *count, int cap, char *names, size_t names_len)
{
    std::lock_guard<std::mutex> lk(g_timer_mu);
    drain_slots();
    size_t pos = 0;
    for (int i = 0; i < KC_COUNT && i < cap; ++i) {
        if (ms) ms[i] = g_timer_ms[i];
        if (count) count[i] = g_timer_cnt[i];
        if (names) {
            size_t l = strlen(kClassNames[i]) + 1;
            if (pos + l <= names_len) { memcpy(names + pos, kClassNames[i], l); pos += l; }
        }
    }
    return KC_COUNT;
}

}  // extern "C"
