// runtime.cpp -- device runtime of libtfhe_mi355x: device selection, the library stream, device
// keys (FFT-domain bootstrapping key + re-laid-out keyswitch key), gate batches and the stage
// entry points of the flat C ABI.  One process drives one GPU; every launch goes on one
// non-blocking stream of default priority, so calls on a device are serialised in issue order.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../hip/engine.h"
#include "internal.h"

using namespace tfhe;

struct tfhe_gpu_key {
    // creation number, unique for the life of the process: a circuit bound to a key remembers it,
    // so a key created later at the same host address is not mistaken for the freed one
    uint64_t id = 0;
    DevParams p{};
    void* bk64 = nullptr;
    void* bk32 = nullptr;
    void* bk32q = nullptr;     // split key in the layout of launch form 6 (N = 1024)
    int32_t* ksk = nullptr;
    int ksk_np = 0;
    int split_bits = 24;
    // scratch of the batch entry points (grown on demand, owned by the key)
    int32_t* d_u = nullptr;
    size_t d_u_elems = 0;
    std::vector<void*> io;        // host-array staging buffers on the device
    std::vector<size_t> io_bytes;
    const TFheGateBootstrappingCloudKeySet* owner = nullptr;
    DevKeyView view() const {
        DevKeyView v;
        v.p = p;
        v.bk64 = bk64;
        v.bk32 = bk32;
        v.bk32q = bk32q;
        v.split_bits = split_bits;
        v.ksk = ksk;
        v.ksk_np = ksk_np;
        return v;
    }
};

namespace {

struct ProfRec {
    std::string name;
    hipEvent_t e0, e1;
};

struct Ctx {
    std::recursive_mutex mu;
    bool inited = false;
    int device = 0;
    hipStream_t stream = nullptr;
    std::map<int, DevTwiddles> tw;
    int split_bits = 0;          // 0: chosen per ring size (split_bits_for)
    int32_t default_prec = TFHE_PREC_FP64;
    bool prof = false;
    std::vector<ProfRec> pending;
    std::map<std::string, std::pair<double, int64_t>> prof_acc;
    // keys created for upstream key sets
    std::map<const TFheGateBootstrappingCloudKeySet*, tfhe_gpu_key*> by_cloud;
    std::vector<const TFheGateBootstrappingCloudKeySet*> known;
};

Ctx& ctx() {
    static Ctx* c = new Ctx();  // never destroyed: device memory is released when the process exits
    return *c;
}

thread_local std::string g_last_error;

int hip_fail(const char* what, hipError_t e) {
    char buf[512];
    snprintf(buf, sizeof buf, "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
    g_last_error = buf;
    return TFHE_E_HIP;
}

int arg_fail(const char* what) {
    g_last_error = what;
    return TFHE_E_ARG;
}

#define HIPR(call)                                            \
    do {                                                      \
        hipError_t _e = (call);                               \
        if (_e != hipSuccess) return hip_fail(#call, _e);     \
    } while (0)

int ensure_init() {
    Ctx& c = ctx();
    if (c.inited) return TFHE_OK;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0) {
        g_last_error = "no HIP device available";
        return TFHE_E_NODEV;
    }
    HIPR(hipSetDevice(c.device));
    HIPR(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    // device address of the fixed scratch of tfhe_gpu_gate_batch_dev_noalloc, with the device current
    HIPR(scratch_init(c.device));
    c.inited = true;
    return TFHE_OK;
}

int ensure_twiddles(int N, DevTwiddles** out) {
    Ctx& c = ctx();
    auto it = c.tw.find(N);
    if (it == c.tw.end()) {
        DevTwiddles t;
        hipError_t e = build_twiddles(t, N);
        if (e != hipSuccess) {
            free_twiddles(t);
            return hip_fail("build_twiddles", e);
        }
        it = c.tw.emplace(N, t).first;
    }
    *out = &it->second;
    return TFHE_OK;
}

// Profiling: a hipEvent pair around each launch on the library stream.
struct ProfScope {
    Ctx& c;
    const char* name;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ProfScope(Ctx& cc, const char* n) : c(cc), name(n) {
        if (!c.prof) return;
        // launches being captured into a graph are not timed (their events would be graph nodes)
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(c.stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) { e0 = e1 = nullptr; return; }
        (void)hipEventRecord(e0, c.stream);
    }
    ~ProfScope() {
        if (!e0) return;
        (void)hipEventRecord(e1, c.stream);
        c.pending.push_back(ProfRec{name, e0, e1});
    }
};

void drain_profile(Ctx& c) {
    for (auto& r : c.pending) {
        (void)hipEventSynchronize(r.e1);
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.e0, r.e1) == hipSuccess) {
            auto& a = c.prof_acc[r.name];
            a.first += ms;
            a.second += 1;
        }
        (void)hipEventDestroy(r.e0);
        (void)hipEventDestroy(r.e1);
    }
    c.pending.clear();
}

int grow(void** p, size_t* have, size_t need_bytes) {
    if (*have >= need_bytes) return TFHE_OK;
    if (*p) HIPR(hipFree(*p));
    *p = nullptr;
    *have = 0;
    HIPR(hipMalloc(p, need_bytes));
    *have = need_bytes;
    return TFHE_OK;
}

int key_u_buffer(tfhe_gpu_key* k, size_t elems) {
    size_t have = k->d_u_elems * sizeof(int32_t);
    void* p = k->d_u;
    int rc = grow(&p, &have, elems * sizeof(int32_t));
    k->d_u = (int32_t*)p;
    k->d_u_elems = have / sizeof(int32_t);
    return rc;
}

int key_io_buffer(tfhe_gpu_key* k, int slot, size_t bytes, void** out) {
    if ((int)k->io.size() <= slot) {
        k->io.resize(slot + 1, nullptr);
        k->io_bytes.resize(slot + 1, 0);
    }
    int rc = grow(&k->io[slot], &k->io_bytes[slot], bytes ? bytes : 4);
    *out = k->io[slot];
    return rc;
}

void gate_coeffs(int gate, int32_t* C, int32_t* al, int32_t* be) {
    const int32_t e8 = (int32_t)(1u << 29), e4 = (int32_t)(1u << 30);  // 1/8, 1/4 of the torus
    switch (gate) {
        case TFHE_GATE_NAND: *C = e8; *al = -1; *be = -1; break;
        case TFHE_GATE_AND: *C = -e8; *al = 1; *be = 1; break;
        case TFHE_GATE_OR: *C = e8; *al = 1; *be = 1; break;          // tb :273-276
        case TFHE_GATE_XOR: *C = e4; *al = 2; *be = 2; break;
        case TFHE_GATE_XNOR: *C = -e4; *al = -2; *be = -2; break;
        case TFHE_GATE_NOR: *C = -e8; *al = -1; *be = -1; break;
        case TFHE_GATE_ANDNY: *C = -e8; *al = -1; *be = 1; break;
        case TFHE_GATE_ANDYN: *C = -e8; *al = 1; *be = -1; break;
        case TFHE_GATE_ORNY: *C = e8; *al = -1; *be = 1; break;
        case TFHE_GATE_ORYN: *C = e8; *al = 1; *be = -1; break;
        default: *C = 0; *al = 0; *be = 0; break;
    }
}

int check_key_prec(tfhe_gpu_key* key, int32_t prec) {
    if (!key) return arg_fail("null device key");
    if (prec == TFHE_PREC_FP64 && !key->bk64) return arg_fail("device key has no FP64 transform");
    if (prec == TFHE_PREC_SPLIT32 && !key->bk32) return arg_fail("device key has no split-FP32 transform");
    if (prec != TFHE_PREC_FP64 && prec != TFHE_PREC_SPLIT32) return arg_fail("unknown precision");
    return TFHE_OK;
}

const char* boot_name(int32_t prec) { return prec == TFHE_PREC_FP64 ? "bootstrap_fp64" : "bootstrap_split32"; }

// device-resident gate batch; all pointers are device pointers.  The extracted samples go to the
// key's scratch (grown with hipMalloc on demand), or, with fixed_u (the no-allocation entry point),
// to that buffer of fixed_elems int32: then nothing is allocated and no profiling events are made.
int gate_batch_dev_impl(tfhe_gpu_key* key, int32_t prec, int32_t gate, int32_t* out, const int32_t* a,
                        const int32_t* b, const int32_t* c, int32_t count, int32_t* fixed_u = nullptr,
                        size_t fixed_elems = 0) {
    Ctx& cx = ctx();
    if (count < 0) return arg_fail("negative count");
    if (count == 0) return TFHE_OK;
    if (gate < TFHE_GATE_NAND || gate > TFHE_GATE_MUX) return arg_fail("unknown gate");
    if (!out || !a || !b || (gate == TFHE_GATE_MUX && !c)) return arg_fail("null ciphertext array");
    int rc = check_key_prec(key, prec);
    if (rc) return rc;
    DevTwiddles* tw;
    if (!fixed_u) {
        if ((rc = ensure_twiddles(key->p.N, &tw))) return rc;
    } else {
        // built when the key was created; looked up only, never built here
        auto it = cx.tw.find(key->p.N);
        if (it == cx.tw.end()) return arg_fail("no transform tables for this ring size (create the key first)");
        tw = &it->second;
    }
    const int N = key->p.N;
    const int32_t MU = (int32_t)(1u << 29);
    const int nboot = (gate == TFHE_GATE_MUX) ? 2 * count : count;
    const size_t u_elems = (size_t)nboot * (N + 1);
    int32_t* u;
    if (fixed_u) {
        if (u_elems > fixed_elems) {
            char msg[256];
            snprintf(msg, sizeof msg,
                     "tfhe_gpu_gate_batch_dev_noalloc: %d gates need %zu int32 of scratch, the fixed scratch holds "
                     "%zu (at most %zu bootstraps at N=%d; a MUX is two)",
                     count, u_elems, fixed_elems, fixed_elems / (N + 1), N);
            return arg_fail(msg);
        }
        u = fixed_u;
    } else {
        if ((rc = key_u_buffer(key, u_elems))) return rc;
        u = key->d_u;
    }
    const bool prof = fixed_u == nullptr;
    BootLaunch L;
    L.mode = 0;
    L.count = nboot;
    L.mu = MU;
    L.u_out = u;
    if (fused_keyswitch() && prec == TFHE_PREC_SPLIT32 && N == 1024 && gate != TFHE_GATE_MUX) {
        // one launch: the bootstrap kernel keyswitches its own sample into the output
        L.nseg = 1;
        L.seg_start[1] = count;
        L.in_a[0] = a;
        L.in_b[0] = b;
        gate_coeffs(gate, &L.cst[0], &L.alpha[0], &L.beta[0]);
        L.ks_out = out;
        std::unique_ptr<ProfScope> ps(prof ? new ProfScope(cx, "bootstrap_split32_fused_keyswitch") : nullptr);
        hipError_t e = launch_bootstrap(key->view(), *tw, (Prec)prec, L, cx.stream);
        if (e != hipSuccess) return hip_fail("fused bootstrap+keyswitch launch", e);
        return TFHE_OK;
    }
    if (gate != TFHE_GATE_MUX) {
        L.nseg = 1;
        L.seg_start[1] = count;
        L.in_a[0] = a;
        L.in_b[0] = b;
        gate_coeffs(gate, &L.cst[0], &L.alpha[0], &L.beta[0]);
    } else {
        // upstream bootsMUX: woKS(AND(a, b)) + woKS(AND(NOT a, c)) + (0, 1/8), then one keyswitch
        L.nseg = 2;
        L.seg_start[1] = count;
        L.seg_start[2] = 2 * count;
        L.in_a[0] = a; L.in_b[0] = b;
        gate_coeffs(TFHE_GATE_AND, &L.cst[0], &L.alpha[0], &L.beta[0]);
        L.in_a[1] = a; L.in_b[1] = c;
        gate_coeffs(TFHE_GATE_ANDNY, &L.cst[1], &L.alpha[1], &L.beta[1]);
    }
    {
        std::unique_ptr<ProfScope> ps(prof ? new ProfScope(cx, boot_name(prec)) : nullptr);
        hipError_t e = launch_bootstrap(key->view(), *tw, (Prec)prec, L, cx.stream);
        if (e != hipSuccess) return hip_fail("bootstrap launch", e);
    }
    {
        std::unique_ptr<ProfScope> ps(prof ? new ProfScope(cx, "keyswitch") : nullptr);
        hipError_t e = (gate == TFHE_GATE_MUX)
                           ? launch_keyswitch(key->view(), out, u, u + (size_t)count * (N + 1), MU, count, cx.stream)
                           : launch_keyswitch(key->view(), out, u, nullptr, 0, count, cx.stream);
        if (e != hipSuccess) return hip_fail("keyswitch launch", e);
    }
    return TFHE_OK;
}

int h2d(void* d, const void* h, size_t bytes) {
    if (!bytes) return TFHE_OK;
    HIPR(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, ctx().stream));
    return TFHE_OK;
}
int d2h_sync(void* h, const void* d, size_t bytes) {
    if (bytes) HIPR(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, ctx().stream));
    HIPR(hipStreamSynchronize(ctx().stream));
    return TFHE_OK;
}

}  // namespace

// =============================================================================================
// internal helpers used by tfhe_api.cpp
// =============================================================================================
namespace tfhe {

void set_error(const std::string& msg) { g_last_error = msg; }

void fatal(const char* where, int code) {
    fprintf(stderr, "tfhe_mi355x: %s failed (%d): %s\n", where, code, g_last_error.c_str());
    abort();
}

int32_t default_precision() { return ctx().default_prec; }

void key_params(const tfhe_gpu_key* key, tfhe_flat_params* out) {
    *out = tfhe_flat_params{key->p.n, key->p.N, 1, key->p.l, key->p.Bgbit, key->p.ks_t, key->p.ks_basebit};
}

tfhe_gpu_key* device_key_for(const TFheGateBootstrappingCloudKeySet* bk) {
    Ctx& c = ctx();
    std::lock_guard<std::recursive_mutex> g(c.mu);
    auto it = c.by_cloud.find(bk);
    if (it != c.by_cloud.end()) return it->second;
    tfhe_gpu_key* key = nullptr;
    if (tfhe_gpu_key_upload(bk, TFHE_PREC_MASK_FP64 | TFHE_PREC_MASK_SPLIT32, &key) != TFHE_OK) return nullptr;
    return key;  // registered by tfhe_gpu_key_upload
}

void register_keyset(const TFheGateBootstrappingCloudKeySet* bk) {
    Ctx& c = ctx();
    std::lock_guard<std::recursive_mutex> g(c.mu);
    c.known.push_back(bk);
}

// The upstream internals receive only a part of a key set (its bkFFT, its TGswSampleFFT array or
// its keyswitch key); the owning cloud key set is found among the ones this library created, and
// its device key is built on first use.
template <class Pred>
static tfhe_gpu_key* device_key_where(Pred pred) {
    Ctx& c = ctx();
    std::lock_guard<std::recursive_mutex> g(c.mu);
    for (const TFheGateBootstrappingCloudKeySet* bk : c.known)
        if (pred(bk)) return device_key_for(bk);
    return nullptr;
}

tfhe_gpu_key* device_key_for_lwebk(const LweBootstrappingKeyFFT* bkfft) {
    return device_key_where([&](const TFheGateBootstrappingCloudKeySet* bk) { return bk->bkFFT == bkfft; });
}

tfhe_gpu_key* device_key_for_bkfft(const TGswSampleFFT* s) {
    return device_key_where(
        [&](const TFheGateBootstrappingCloudKeySet* bk) { return bk->bkFFT && bk->bkFFT->bkFFT == s; });
}

tfhe_gpu_key* device_key_for_ksk(const LweKeySwitchKey* ks) {
    return device_key_where([&](const TFheGateBootstrappingCloudKeySet* bk) { return bk->bk && bk->bk->ks == ks; });
}

void forget_device_keys(const TFheGateBootstrappingCloudKeySet* bk) {
    Ctx& c = ctx();
    std::lock_guard<std::recursive_mutex> g(c.mu);
    for (size_t i = 0; i < c.known.size(); ++i)
        if (c.known[i] == bk) { c.known.erase(c.known.begin() + i); break; }
    auto it = c.by_cloud.find(bk);
    if (it == c.by_cloud.end()) return;
    tfhe_gpu_key* k = it->second;
    c.by_cloud.erase(it);
    tfhe_gpu_key_free(k);
}

std::recursive_mutex& lib_mutex() { return ctx().mu; }
int lib_ensure_init() { return ensure_init(); }
hipStream_t lib_stream() { return ctx().stream; }
int lib_twiddles(int N, DevTwiddles** out) { return ensure_twiddles(N, out); }
int lib_hip_fail(const char* what, hipError_t e) { return hip_fail(what, e); }
int lib_arg_fail(const char* what) { return arg_fail(what); }
int key_check_prec(tfhe_gpu_key* key, int32_t prec) { return check_key_prec(key, prec); }
DevKeyView key_view(const tfhe_gpu_key* key) { return key->view(); }
uint64_t key_id(const tfhe_gpu_key* key) { return key->id; }
void gate_coefficients(int gate, int32_t* C, int32_t* alpha, int32_t* beta) { gate_coeffs(gate, C, alpha, beta); }

LibProf::LibProf(const char* name) : impl(new ProfScope(ctx(), name)) {}
LibProf::~LibProf() { delete static_cast<ProfScope*>(impl); }

}  // namespace tfhe

// =============================================================================================
// flat C ABI
// =============================================================================================
extern "C" const char* tfhe_gpu_last_error(void) { return g_last_error.c_str(); }

extern "C" int tfhe_gpu_init(int32_t device) {
    Ctx& c = ctx();
    std::lock_guard<std::recursive_mutex> g(c.mu);
    if (c.inited) return device == c.device ? TFHE_OK : arg_fail("library already bound to another device");
    c.device = device;
    return ensure_init();
}

extern "C" int tfhe_gpu_device_info(char* name, int32_t name_len, int32_t* cu_count, int64_t* hbm_bytes) {
    Ctx& c = ctx();
    std::lock_guard<std::recursive_mutex> g(c.mu);
    int rc = ensure_init();
    if (rc) return rc;
    hipDeviceProp_t prop;
    HIPR(hipGetDeviceProperties(&prop, c.device));
    if (name && name_len > 0) {
        snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (cu_count) *cu_count = prop.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = (int64_t)prop.totalGlobalMem;
    return TFHE_OK;
}

extern "C" int tfhe_gpu_sync(void) {
    Ctx& c = ctx();
    std::lock_guard<std::recursive_mutex> g(c.mu);
    int rc = ensure_init();
    if (rc) return rc;
    HIPR(hipStreamSynchronize(c.stream));
    return TFHE_OK;
}

extern "C" int tfhe_gpu_set_split_bits(int32_t bits) {
    if (bits != 0 && (bits < 8 || bits > 28)) return arg_fail("split bits must be 0 (default) or in [8, 28]");
    ctx().split_bits = bits;
    return TFHE_OK;
}

extern "C" int tfhe_gpu_set_fp64_shared_rows(int32_t on) {
    if (on < 0 || on > 2) return arg_fail("fp64 shared-rows flag must be 0, 1 or 2");
    set_boot64_shared(on);
    return TFHE_OK;
}

extern "C" int tfhe_gpu_set_split_mac(int32_t mode) {
    if (mode != 0 && mode != 1) return arg_fail("split MAC form must be 0 (plain) or 1 (compensated)");
    std::lock_guard<std::recursive_mutex> g(ctx().mu);
    set_split_mac(mode);
    return TFHE_OK;
}

extern "C" int tfhe_gpu_set_fused_keyswitch(int32_t on) {
    if (on != 0 && on != 1) return arg_fail("fused keyswitch flag must be 0 or 1");
    std::lock_guard<std::recursive_mutex> g(ctx().mu);
    set_fused_keyswitch(on);
    return TFHE_OK;
}

extern "C" int tfhe_gpu_set_launch_form(int32_t form) {
    if (form < 0 || form > 6) return arg_fail("launch form must be in [0, 6]");
    std::lock_guard<std::recursive_mutex> g(ctx().mu);
    set_forced_form(form);
    return TFHE_OK;
}

extern "C" int tfhe_gpu_set_default_precision(int32_t prec) {
    if (prec != TFHE_PREC_FP64 && prec != TFHE_PREC_SPLIT32) return arg_fail("unknown precision");
    ctx().default_prec = prec;
    return TFHE_OK;
}

// Split point of the FP32 key halves when none was set.  The hi product (digits x Khi summed over
// (k+1) l N terms) must stay below ~2^21 for FP32 to round it to the exact integer: with
// Khi < 2^(31-s), Bg/2 = 64 and N = 1024 that holds for s = 24; N = 2048 with Bg/2 = 128 grows the
// sum by ~2^2, so its hi half is made 2 bits narrower.
static int split_bits_for(const DevParams& p) { return p.N <= 1024 ? 24 : 25; }

extern "C" int tfhe_gpu_key_create(const tfhe_flat_params* fp, const int32_t* bk, const int32_t* ksk,
                                   int32_t prec_mask, tfhe_gpu_key** out) {
    Ctx& c = ctx();
    std::lock_guard<std::recursive_mutex> g(c.mu);
    if (!fp || !bk || !ksk || !out) return arg_fail("null argument to tfhe_gpu_key_create");
    if ((prec_mask & 3) == 0) return arg_fail("empty precision mask");
    DevParams p{fp->n, fp->N, fp->l, fp->Bgbit, fp->ks_t, fp->ks_basebit};
    if (fp->k != 1 || !params_supported(p)) {
        g_last_error = "unsupported parameter set (device path: k=1, (N,l,Bgbit) in {(1024,3,7),(2048,3,8)}, "
                       "ks basebit 2, ks t in {8,10}, n <= 1024)";
        return TFHE_E_UNSUPPORTED;
    }
    int rc = ensure_init();
    if (rc) return rc;
    DevTwiddles* tw;
    if ((rc = ensure_twiddles(p.N, &tw))) return rc;
    tfhe_gpu_key* k = new tfhe_gpu_key();
    static uint64_t next_id = 0;
    k->id = ++next_id;   // under the library mutex
    k->p = p;
    k->split_bits = c.split_bits ? c.split_bits : split_bits_for(p);
    const int M = p.N / 2, kpl = 2 * p.l;
    const size_t npoly = (size_t)p.n * kpl * 2;
    const size_t bk_time_bytes = npoly * p.N * sizeof(int32_t);
    const size_t bk_fft_bytes = npoly * M * 16;
    const int base = 1 << p.ks_basebit;
    const size_t ksk_in_bytes = (size_t)p.N * p.ks_t * base * (p.n + 1) * sizeof(int32_t);
    k->ksk_np = ((p.n + 1) + 63) / 64 * 64;
    const size_t ksk_bytes = ksk_device_bytes(p, k->ksk_np);
    int32_t* d_bk = nullptr;
    int32_t* d_ksk_in = nullptr;
    auto cleanup_fail = [&](int code) {
        if (d_bk) (void)hipFree(d_bk);
        if (d_ksk_in) (void)hipFree(d_ksk_in);
        tfhe_gpu_key_free(k);
        return code;
    };
    hipError_t e;
    if ((e = hipMalloc(&d_bk, bk_time_bytes)) != hipSuccess) return cleanup_fail(hip_fail("hipMalloc bk", e));
    if ((e = hipMemcpyAsync(d_bk, bk, bk_time_bytes, hipMemcpyHostToDevice, c.stream)) != hipSuccess)
        return cleanup_fail(hip_fail("copy bk", e));
    if (prec_mask & TFHE_PREC_MASK_FP64)
        if ((e = hipMalloc(&k->bk64, bk_fft_bytes)) != hipSuccess) return cleanup_fail(hip_fail("hipMalloc bk64", e));
    if (prec_mask & TFHE_PREC_MASK_SPLIT32)
        if ((e = hipMalloc(&k->bk32, bk_fft_bytes)) != hipSuccess) return cleanup_fail(hip_fail("hipMalloc bk32", e));
    if ((e = launch_keyprep(p, *tw, d_bk, k->bk64, k->bk32, k->split_bits, c.stream)) != hipSuccess)
        return cleanup_fail(hip_fail("keyprep", e));
    if (k->bk32 && p.N == 1024 && forced_form() == 6) {
        // the same split key arranged for the matrix-core form (4 bootstraps per wavefront), made
        // only for keys created while that form is selected (it is not the default: 124 MB more)
        if ((e = hipMalloc(&k->bk32q, (size_t)p.n * 6 * 32 * 64 * 16)) != hipSuccess)
            return cleanup_fail(hip_fail("hipMalloc bk32q", e));
        if ((e = launch_keyprep_q4(p, k->bk32, k->bk32q, c.stream)) != hipSuccess)
            return cleanup_fail(hip_fail("keyprep q4", e));
    }
    if ((e = hipMalloc(&d_ksk_in, ksk_in_bytes)) != hipSuccess) return cleanup_fail(hip_fail("hipMalloc ksk", e));
    if ((e = hipMemcpyAsync(d_ksk_in, ksk, ksk_in_bytes, hipMemcpyHostToDevice, c.stream)) != hipSuccess)
        return cleanup_fail(hip_fail("copy ksk", e));
    if ((e = hipMalloc(&k->ksk, ksk_bytes)) != hipSuccess) return cleanup_fail(hip_fail("hipMalloc ksk dev", e));
    if ((e = launch_ksk_relayout(p, d_ksk_in, k->ksk, k->ksk_np, c.stream)) != hipSuccess)
        return cleanup_fail(hip_fail("ksk relayout", e));
    if ((e = hipStreamSynchronize(c.stream)) != hipSuccess) return cleanup_fail(hip_fail("key sync", e));
    (void)hipFree(d_bk);
    (void)hipFree(d_ksk_in);
    *out = k;
    return TFHE_OK;
}

extern "C" int tfhe_gpu_key_upload(const TFheGateBootstrappingCloudKeySet* bk, int32_t prec_mask, tfhe_gpu_key** out) {
    Ctx& c = ctx();
    std::lock_guard<std::recursive_mutex> g(c.mu);
    if (!bk || !bk->params || !out) return arg_fail("null cloud key");
    const bool fft_only = !bk->bk;
    if (fft_only && (!bk->bkFFT || !bk->bkFFT->bkFFT || !bk->bkFFT->ks))
        return arg_fail("cloud key set without a bootstrapping key");
    tfhe_flat_params fp;
    flat_params_of(bk->params, &fp);
    const size_t bk_elems = (size_t)fp.n * (fp.k + 1) * fp.l * (fp.k + 1) * fp.N;
    const size_t ksk_elems = (size_t)fp.k * fp.N * fp.ks_t * (1 << fp.ks_basebit) * (fp.n + 1);
    std::vector<int32_t> hb(bk_elems), hk(ksk_elems);
    if (fft_only) {
        const int rc = bk_from_lagrange_struct(bk->bkFFT, &fp, hb.data());
        if (rc) return rc;
        flatten_ksk(bk->bkFFT->ks, hk.data());
    } else {
        flatten_bk(bk->bk, hb.data());
        flatten_ksk(bk->bk->ks, hk.data());
    }
    tfhe_gpu_key* k = nullptr;
    int rc = tfhe_gpu_key_create(&fp, hb.data(), hk.data(), prec_mask, &k);
    if (rc) return rc;
    auto it = c.by_cloud.find(bk);
    if (it != c.by_cloud.end()) {
        tfhe_gpu_key* old = it->second;
        it->second = k;
        tfhe_gpu_key_free(old);
    } else {
        c.by_cloud[bk] = k;
    }
    k->owner = bk;
    *out = k;
    return TFHE_OK;
}

extern "C" int tfhe_gpu_key_free(tfhe_gpu_key* k) {
    if (!k) return TFHE_OK;
    Ctx& c = ctx();
    std::lock_guard<std::recursive_mutex> g(c.mu);
    if (c.inited) (void)hipStreamSynchronize(c.stream);
    if (k->owner) {
        auto it = c.by_cloud.find(k->owner);
        if (it != c.by_cloud.end() && it->second == k) c.by_cloud.erase(it);
    }
    void* ps[] = {k->bk64, k->bk32, k->bk32q, k->ksk, k->d_u};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    for (void* p : k->io)
        if (p) (void)hipFree(p);
    delete k;
    return TFHE_OK;
}

extern "C" int tfhe_gpu_key_params(const tfhe_gpu_key* key, tfhe_flat_params* out) {
    if (!key || !out) return arg_fail("null argument");
    key_params(key, out);
    return TFHE_OK;
}

extern "C" int tfhe_gpu_gate_batch_dev(tfhe_gpu_key* key, int32_t prec, int32_t gate, int32_t* d_out,
                                       const int32_t* d_a, const int32_t* d_b, const int32_t* d_c, int32_t count) {
    Ctx& c = ctx();
    std::lock_guard<std::recursive_mutex> g(c.mu);
    int rc = ensure_init();
    if (rc) return rc;
    return gate_batch_dev_impl(key, prec, gate, d_out, d_a, d_b, d_c, count);
}

// The same without allocating: the extracted samples use the library's fixed device scratch (its
// address resolved at init), so no hipMalloc, hipFree or event creation happens here.  Calls that
// overlap on one device share that scratch and race; calls on the library stream are serialised.
extern "C" int tfhe_gpu_gate_batch_dev_noalloc(tfhe_gpu_key* key, int32_t prec, int32_t gate, int32_t* d_out,
                                               const int32_t* d_a, const int32_t* d_b, const int32_t* d_c,
                                               int32_t count) {
    Ctx& c = ctx();
    std::lock_guard<std::recursive_mutex> g(c.mu);
    if (!c.inited) {
        g_last_error = "tfhe_gpu_gate_batch_dev_noalloc: call tfhe_gpu_init (or create the key) first";
        return TFHE_E_NODEV;
    }
    int32_t* u = scratch_ptr(c.device);
    if (!u) {
        g_last_error = "tfhe_gpu_gate_batch_dev_noalloc: fixed scratch not resolved on this device";
        return TFHE_E_NODEV;
    }
    return gate_batch_dev_impl(key, prec, gate, d_out, d_a, d_b, d_c, count, u, kScratchElems);
}

extern "C" int tfhe_gpu_gate_batch(tfhe_gpu_key* key, int32_t prec, int32_t gate, int32_t* out, const int32_t* a,
                                   const int32_t* b, const int32_t* c, int32_t count) {
    Ctx& cx = ctx();
    std::lock_guard<std::recursive_mutex> g(cx.mu);
    int rc = ensure_init();
    if (rc) return rc;
    if (count < 0) return arg_fail("negative count");
    if (count == 0) return TFHE_OK;
    if (!key) return arg_fail("null device key");
    if (!out || !a || !b || (gate == TFHE_GATE_MUX && !c)) return arg_fail("null ciphertext array");
    const size_t bytes = (size_t)count * (key->p.n + 1) * sizeof(int32_t);
    void *da, *db, *dc = nullptr, *dout;
    if ((rc = key_io_buffer(key, 0, bytes, &da)) || (rc = key_io_buffer(key, 1, bytes, &db)) ||
        (rc = key_io_buffer(key, 3, bytes, &dout)))
        return rc;
    if (gate == TFHE_GATE_MUX && (rc = key_io_buffer(key, 2, bytes, &dc))) return rc;
    if ((rc = h2d(da, a, bytes)) || (rc = h2d(db, b, bytes))) return rc;
    if (dc && (rc = h2d(dc, c, bytes))) return rc;
    if ((rc = gate_batch_dev_impl(key, prec, gate, (int32_t*)dout, (const int32_t*)da, (const int32_t*)db,
                                  (const int32_t*)dc, count)))
        return rc;
    return d2h_sync(out, dout, bytes);
}

extern "C" int tfhe_gpu_boots_batch(int32_t gate, LweSample* out, const LweSample* a, const LweSample* b,
                                    const LweSample* c, int32_t count, const TFheGateBootstrappingCloudKeySet* bk) {
    if (count < 0) return arg_fail("negative count");
    if (count == 0) return TFHE_OK;
    if (!out || !a || !b || !bk || (gate == TFHE_GATE_MUX && !c)) return arg_fail("null argument");
    tfhe_gpu_key* key = device_key_for(bk);
    if (!key) return TFHE_E_HIP;
    const int n = key->p.n;
    std::vector<int32_t> fa((size_t)count * (n + 1)), fb(fa.size()), fc, fo(fa.size());
    for (int i = 0; i < count; ++i) {
        memcpy(&fa[(size_t)i * (n + 1)], a[i].a, sizeof(int32_t) * n);
        fa[(size_t)i * (n + 1) + n] = a[i].b;
        memcpy(&fb[(size_t)i * (n + 1)], b[i].a, sizeof(int32_t) * n);
        fb[(size_t)i * (n + 1) + n] = b[i].b;
    }
    if (gate == TFHE_GATE_MUX) {
        fc.resize(fa.size());
        for (int i = 0; i < count; ++i) {
            memcpy(&fc[(size_t)i * (n + 1)], c[i].a, sizeof(int32_t) * n);
            fc[(size_t)i * (n + 1) + n] = c[i].b;
        }
    }
    int rc = tfhe_gpu_gate_batch(key, default_precision(), gate, fo.data(), fa.data(), fb.data(),
                                 fc.empty() ? nullptr : fc.data(), count);
    if (rc) return rc;
    for (int i = 0; i < count; ++i) {
        memcpy(out[i].a, &fo[(size_t)i * (n + 1)], sizeof(int32_t) * n);
        out[i].b = fo[(size_t)i * (n + 1) + n];
        out[i].current_variance = 0.;
    }
    return TFHE_OK;
}

// ---- stage entry points (host arrays) ----

static int boot_stage(tfhe_gpu_key* key, int32_t prec, int mode, int32_t* io, const int32_t* x, const int32_t* bara,
                      const int32_t* row_set, double* raw, int32_t mu, int32_t count, int32_t* u_host,
                      double* raw_hi = nullptr) {
    Ctx& cx = ctx();
    int rc = ensure_init();
    if (rc) return rc;
    if (count < 0) return arg_fail("negative count");
    if (count == 0) return TFHE_OK;
    if ((rc = check_key_prec(key, prec))) return rc;
    DevTwiddles* tw;
    if ((rc = ensure_twiddles(key->p.N, &tw))) return rc;
    const int N = key->p.N, n = key->p.n;
    BootLaunch L;
    L.mode = mode;
    L.count = count;
    L.mu = mu;
    void *dx = nullptr, *dacc = nullptr, *dbara = nullptr, *drow = nullptr, *draw = nullptr, *draw_hi = nullptr;
    const size_t acc_bytes = (size_t)count * 2 * N * sizeof(int32_t);
    if (mode == 0) {
        const size_t xb = (size_t)count * (n + 1) * sizeof(int32_t);
        if ((rc = key_io_buffer(key, 0, xb, &dx)) || (rc = h2d(dx, x, xb))) return rc;
        if ((rc = key_u_buffer(key, (size_t)count * (N + 1)))) return rc;
        L.in_a[0] = (const int32_t*)dx;
        L.alpha[0] = 1;
        L.seg_start[1] = count;
        L.u_out = key->d_u;
        if (io) {
            if ((rc = key_io_buffer(key, 1, acc_bytes, &dacc))) return rc;
            L.acc_io = (int32_t*)dacc;
        }
    } else {
        if ((rc = key_io_buffer(key, 1, acc_bytes, &dacc)) || (rc = h2d(dacc, io, acc_bytes))) return rc;
        L.acc_io = (int32_t*)dacc;
        if (mode == 1) {
            const size_t bb = (size_t)count * n * sizeof(int32_t);
            for (int64_t i = 0; i < (int64_t)count * n; ++i)
                if (bara[i] < 0 || bara[i] >= 2 * N) return arg_fail("bara entries must lie in [0, 2N)");
            if ((rc = key_io_buffer(key, 2, bb, &dbara)) || (rc = h2d(dbara, bara, bb))) return rc;
            L.bara_in = (const int32_t*)dbara;
        } else {
            const size_t rb = (size_t)count * sizeof(int32_t);
            for (int i = 0; i < count; ++i)
                if (row_set[i] < 0 || row_set[i] >= n) return arg_fail("row_set entries must lie in [0, n)");
            if ((rc = key_io_buffer(key, 2, rb, &drow)) || (rc = h2d(drow, row_set, rb))) return rc;
            L.row_set = (const int32_t*)drow;
            if (raw) {
                if ((rc = key_io_buffer(key, 4, acc_bytes * 2, &draw))) return rc;
                L.raw_out = (double*)draw;
            }
            if (raw_hi) {
                if ((rc = key_io_buffer(key, 5, acc_bytes * 2, &draw_hi))) return rc;
                L.raw_hi = (double*)draw_hi;
            }
        }
    }
    {
        ProfScope ps(cx, boot_name(prec));
        hipError_t e = launch_bootstrap(key->view(), *tw, (Prec)prec, L, cx.stream);
        if (e != hipSuccess) return hip_fail("bootstrap launch", e);
    }
    if (mode == 0) {
        if (u_host && (rc = d2h_sync(u_host, key->d_u, (size_t)count * (N + 1) * sizeof(int32_t)))) return rc;
        if (io && (rc = d2h_sync(io, dacc, acc_bytes))) return rc;
    } else {
        if ((rc = d2h_sync(io, dacc, acc_bytes))) return rc;
        if (raw && (rc = d2h_sync(raw, draw, acc_bytes * 2))) return rc;
        if (raw_hi && (rc = d2h_sync(raw_hi, draw_hi, acc_bytes * 2))) return rc;
    }
    return TFHE_OK;
}

extern "C" int tfhe_gpu_bootstrap_woks_batch(tfhe_gpu_key* key, int32_t prec, int32_t* u, const int32_t* x,
                                             int32_t mu, int32_t count) {
    std::lock_guard<std::recursive_mutex> g(ctx().mu);
    if (count > 0 && (!u || !x)) return arg_fail("null argument");
    return boot_stage(key, prec, 0, nullptr, x, nullptr, nullptr, nullptr, mu, count, u);
}

// bootstrap with its accumulator exported (tests): acc [count][2][N]
extern "C" int tfhe_gpu_bootstrap_woks_acc_batch(tfhe_gpu_key* key, int32_t prec, int32_t* u, int32_t* acc,
                                                 const int32_t* x, int32_t mu, int32_t count) {
    std::lock_guard<std::recursive_mutex> g(ctx().mu);
    if (count > 0 && (!u || !x || !acc)) return arg_fail("null argument");
    return boot_stage(key, prec, 0, acc, x, nullptr, nullptr, nullptr, mu, count, u);
}

extern "C" int tfhe_gpu_bootstrap_batch(tfhe_gpu_key* key, int32_t prec, int32_t* out, const int32_t* x, int32_t mu,
                                        int32_t count) {
    Ctx& cx = ctx();
    std::lock_guard<std::recursive_mutex> g(cx.mu);
    if (count < 0) return arg_fail("negative count");
    if (count == 0) return TFHE_OK;
    if (!out || !x) return arg_fail("null argument");
    int rc = boot_stage(key, prec, 0, nullptr, x, nullptr, nullptr, nullptr, mu, count, nullptr);
    if (rc) return rc;
    const size_t ob = (size_t)count * (key->p.n + 1) * sizeof(int32_t);
    void* dout;
    if ((rc = key_io_buffer(key, 3, ob, &dout))) return rc;
    {
        ProfScope ps(cx, "keyswitch");
        hipError_t e = launch_keyswitch(key->view(), (int32_t*)dout, key->d_u, nullptr, 0, count, cx.stream);
        if (e != hipSuccess) return hip_fail("keyswitch launch", e);
    }
    return d2h_sync(out, dout, ob);
}

extern "C" int tfhe_gpu_blind_rotate_batch(tfhe_gpu_key* key, int32_t prec, int32_t* acc, const int32_t* bara,
                                           int32_t count) {
    std::lock_guard<std::recursive_mutex> g(ctx().mu);
    if (count > 0 && (!acc || !bara)) return arg_fail("null argument");
    return boot_stage(key, prec, 1, acc, nullptr, bara, nullptr, nullptr, 0, count, nullptr);
}

extern "C" int tfhe_gpu_extern_product_batch(tfhe_gpu_key* key, int32_t prec, int32_t* acc, const int32_t* row_set,
                                             double* raw, int32_t count) {
    std::lock_guard<std::recursive_mutex> g(ctx().mu);
    if (count > 0 && (!acc || !row_set)) return arg_fail("null argument");
    return boot_stage(key, prec, 2, acc, nullptr, nullptr, row_set, raw, 0, count, nullptr);
}

extern "C" int tfhe_gpu_extern_product_split_hi_batch(tfhe_gpu_key* key, int32_t* acc, const int32_t* row_set,
                                                      double* raw_hi, int32_t count) {
    std::lock_guard<std::recursive_mutex> g(ctx().mu);
    if (count > 0 && (!acc || !row_set || !raw_hi)) return arg_fail("null argument");
    return boot_stage(key, TFHE_PREC_SPLIT32, 2, acc, nullptr, nullptr, row_set, nullptr, 0, count, nullptr, raw_hi);
}

extern "C" int tfhe_gpu_key_split_bits(const tfhe_gpu_key* key, int32_t* bits) {
    if (!key || !bits) return arg_fail("null argument");
    *bits = key->split_bits;
    return TFHE_OK;
}

extern "C" int tfhe_gpu_keyswitch_batch(tfhe_gpu_key* key, int32_t* out, const int32_t* u, int32_t count) {
    Ctx& cx = ctx();
    std::lock_guard<std::recursive_mutex> g(cx.mu);
    int rc = ensure_init();
    if (rc) return rc;
    if (count < 0) return arg_fail("negative count");
    if (count == 0) return TFHE_OK;
    if (!key || !out || !u) return arg_fail("null argument");
    const int N = key->p.N, n = key->p.n;
    const size_t ub = (size_t)count * (N + 1) * sizeof(int32_t), ob = (size_t)count * (n + 1) * sizeof(int32_t);
    void *du, *dout;
    if ((rc = key_io_buffer(key, 0, ub, &du)) || (rc = key_io_buffer(key, 3, ob, &dout)) || (rc = h2d(du, u, ub)))
        return rc;
    {
        ProfScope ps(cx, "keyswitch");
        hipError_t e = launch_keyswitch(key->view(), (int32_t*)dout, (const int32_t*)du, nullptr, 0, count, cx.stream);
        if (e != hipSuccess) return hip_fail("keyswitch launch", e);
    }
    return d2h_sync(out, dout, ob);
}

namespace {
// scratch for stage calls that have no key
struct Scratch {
    std::vector<void*> p;
    std::vector<size_t> b;
    int get(int slot, size_t bytes, void** out) {
        if ((int)p.size() <= slot) { p.resize(slot + 1, nullptr); b.resize(slot + 1, 0); }
        int rc = grow(&p[slot], &b[slot], bytes ? bytes : 4);
        *out = p[slot];
        return rc;
    }
};
Scratch& scratch() {
    static Scratch* s = new Scratch();
    return *s;
}
}  // namespace

extern "C" int tfhe_gpu_extract_batch(const tfhe_flat_params* p, int32_t* u, const int32_t* acc, int32_t count) {
    Ctx& cx = ctx();
    std::lock_guard<std::recursive_mutex> g(cx.mu);
    int rc = ensure_init();
    if (rc) return rc;
    if (!p || count < 0) return arg_fail("bad argument");
    if (count == 0) return TFHE_OK;
    if (!u || !acc) return arg_fail("null argument");
    if (p->k != 1 || (p->N != 1024 && p->N != 2048)) return arg_fail("unsupported N/k");
    const int N = p->N;
    const size_t ab = (size_t)count * 2 * N * 4, ub = (size_t)count * (N + 1) * 4;
    void *da, *du;
    if ((rc = scratch().get(0, ab, &da)) || (rc = scratch().get(1, ub, &du)) || (rc = h2d(da, acc, ab))) return rc;
    {
        ProfScope ps(cx, "extract");
        hipError_t e = launch_extract(N, (int32_t*)du, (const int32_t*)da, count, cx.stream);
        if (e != hipSuccess) return hip_fail("extract launch", e);
    }
    return d2h_sync(u, du, ub);
}

extern "C" int tfhe_gpu_rotate_decompose_batch(const tfhe_flat_params* p, int32_t* rot, int32_t* digits,
                                               const int32_t* acc, const int32_t* bara, int32_t count) {
    Ctx& cx = ctx();
    std::lock_guard<std::recursive_mutex> g(cx.mu);
    int rc = ensure_init();
    if (rc) return rc;
    if (!p || count < 0) return arg_fail("bad argument");
    if (count == 0) return TFHE_OK;
    if (!rot || !digits || !acc || !bara) return arg_fail("null argument");
    DevParams dp{p->n, p->N, p->l, p->Bgbit, p->ks_t, p->ks_basebit};
    if (p->k != 1 || !params_supported(dp)) return arg_fail("unsupported parameter set");
    const int N = p->N;
    for (int i = 0; i < count; ++i)
        if (bara[i] < 0 || bara[i] >= 2 * N) return arg_fail("bara entries must lie in [0, 2N)");
    const size_t ab = (size_t)count * 2 * N * 4, db = (size_t)count * 2 * p->l * N * 4, bb = (size_t)count * 4;
    void *da, *dr, *dd, *dbara;
    if ((rc = scratch().get(0, ab, &da)) || (rc = scratch().get(1, ab, &dr)) || (rc = scratch().get(2, db, &dd)) ||
        (rc = scratch().get(3, bb, &dbara)) || (rc = h2d(da, acc, ab)) || (rc = h2d(dbara, bara, bb)))
        return rc;
    hipError_t e = launch_rotate_decompose(dp, (int32_t*)dr, (int32_t*)dd, (const int32_t*)da, (const int32_t*)dbara,
                                           count, cx.stream);
    if (e != hipSuccess) return hip_fail("rotate_decompose launch", e);
    if ((rc = d2h_sync(rot, dr, ab))) return rc;
    return d2h_sync(digits, dd, db);
}

static int fft_stage(int32_t N, int32_t prec, int inverse, double* out, const void* in, int32_t count) {
    Ctx& cx = ctx();
    std::lock_guard<std::recursive_mutex> g(cx.mu);
    int rc = ensure_init();
    if (rc) return rc;
    if (count < 0) return arg_fail("negative count");
    if (count == 0) return TFHE_OK;
    if (!out || !in) return arg_fail("null argument");
    if (N != 1024 && N != 2048) return arg_fail("N must be 1024 or 2048");
    if (prec != TFHE_PREC_FP64 && prec != TFHE_PREC_SPLIT32) return arg_fail("unknown precision");
    DevTwiddles* tw;
    if ((rc = ensure_twiddles(N, &tw))) return rc;
    if (inverse == 2) {
        // the matrix-core form takes gadget-digit inputs, exact in FP16
        if (prec != TFHE_PREC_SPLIT32) return arg_fail("digit transform: split FP32 only");
        const int32_t* p = (const int32_t*)in;
        for (size_t i = 0; i < (size_t)count * N; ++i)
            if (p[i] <= -2048 || p[i] >= 2048) return arg_fail("digit transform: inputs must lie in (-2048, 2048)");
    }
    const size_t ib = (size_t)count * N * (inverse == 1 ? 8 : 4), ob = (size_t)count * N * 8;
    void *di, *dout;
    if ((rc = scratch().get(0, ib, &di)) || (rc = scratch().get(1, ob, &dout)) || (rc = h2d(di, in, ib))) return rc;
    hipError_t e = launch_fft_test(*tw, N, (Prec)prec, inverse, (double*)dout, di, count, cx.stream);
    if (e != hipSuccess) return hip_fail("fft launch", e);
    return d2h_sync(out, dout, ob);
}

extern "C" int tfhe_gpu_fft_q4_batch(int32_t inverse, double* out, const void* in, int32_t count) {
    Ctx& cx = ctx();
    std::lock_guard<std::recursive_mutex> g(cx.mu);
    int rc = ensure_init();
    if (rc) return rc;
    if (count < 0) return arg_fail("negative count");
    if (count == 0) return TFHE_OK;
    if (!out || !in) return arg_fail("null argument");
    DevTwiddles* tw;
    if ((rc = ensure_twiddles(1024, &tw))) return rc;
    const size_t ib = (size_t)count * 1024 * (inverse ? 8 : 4), ob = (size_t)count * 1024 * 8;
    void *di, *dout;
    if ((rc = scratch().get(0, ib, &di)) || (rc = scratch().get(1, ob, &dout)) || (rc = h2d(di, in, ib))) return rc;
    hipError_t e = launch_fft_test_q4(*tw, inverse, (double*)dout, di, count, cx.stream);
    if (e != hipSuccess) return hip_fail("fft q4 launch", e);
    return d2h_sync(out, dout, ob);
}

extern "C" int tfhe_gpu_fft_forward_batch(int32_t N, int32_t prec, double* out, const int32_t* in, int32_t count) {
    return fft_stage(N, prec, 0, out, in, count);
}
extern "C" int tfhe_gpu_fft_inverse_batch(int32_t N, int32_t prec, double* out, const double* in, int32_t count) {
    return fft_stage(N, prec, 1, out, in, count);
}
extern "C" int tfhe_gpu_fft_forward_digits_batch(int32_t N, double* out, const int32_t* in, int32_t count) {
    return fft_stage(N, TFHE_PREC_SPLIT32, 2, out, in, count);
}
extern "C" int tfhe_gpu_split_digits_batch(int32_t Bgbit, int32_t variant, float* out, const int32_t* x,
                                           int32_t count) {
    Ctx& cx = ctx();
    std::lock_guard<std::recursive_mutex> g(cx.mu);
    int rc = ensure_init();
    if (rc) return rc;
    if (count < 0) return arg_fail("negative count");
    if (count == 0) return TFHE_OK;
    if (!out || !x) return arg_fail("null argument");
    if (Bgbit != 7 && Bgbit != 8) return arg_fail("Bgbit must be 7 or 8 (l = 3)");
    if (variant != 0 && variant != 1) return arg_fail("variant must be 0 or 1");
    const size_t ib = (size_t)count * 2 * 4, ob = (size_t)count * 6 * 4;
    void *di, *dout;
    if ((rc = scratch().get(0, ib, &di)) || (rc = scratch().get(1, ob, &dout)) || (rc = h2d(di, x, ib))) return rc;
    hipError_t e = launch_digits_test(Bgbit, variant, (float*)dout, (const int32_t*)di, count, cx.stream);
    if (e != hipSuccess) return hip_fail("digits launch", e);
    return d2h_sync(out, dout, ob);
}

// ---- device memory helpers ----
extern "C" int tfhe_gpu_malloc(void** ptr, size_t bytes) {
    std::lock_guard<std::recursive_mutex> g(ctx().mu);
    int rc = ensure_init();
    if (rc) return rc;
    if (!ptr) return arg_fail("null argument");
    HIPR(hipMalloc(ptr, bytes ? bytes : 4));
    return TFHE_OK;
}
extern "C" int tfhe_gpu_free(void* ptr) {
    std::lock_guard<std::recursive_mutex> g(ctx().mu);
    if (!ptr) return TFHE_OK;
    HIPR(hipStreamSynchronize(ctx().stream));
    HIPR(hipFree(ptr));
    return TFHE_OK;
}
extern "C" int tfhe_gpu_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    std::lock_guard<std::recursive_mutex> g(ctx().mu);
    int rc = ensure_init();
    if (rc) return rc;
    if ((rc = h2d(dst, src, bytes))) return rc;
    HIPR(hipStreamSynchronize(ctx().stream));
    return TFHE_OK;
}
extern "C" int tfhe_gpu_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    std::lock_guard<std::recursive_mutex> g(ctx().mu);
    int rc = ensure_init();
    if (rc) return rc;
    return d2h_sync(dst, src, bytes);
}

extern "C" int tfhe_gpu_decrypt_batch_dev(int32_t* bits, const int32_t* d_samples, const int32_t* lwe_key, int32_t n,
                                          int32_t count) {
    Ctx& cx = ctx();
    std::lock_guard<std::recursive_mutex> g(cx.mu);
    int rc = ensure_init();
    if (rc) return rc;
    if (count < 0 || n <= 0) return arg_fail("bad size");
    if (count == 0) return TFHE_OK;
    if (!bits || !d_samples || !lwe_key) return arg_fail("null argument");
    void *dk, *db;
    if ((rc = scratch().get(4, (size_t)n * 4, &dk)) || (rc = scratch().get(5, (size_t)count * 4, &db)) ||
        (rc = h2d(dk, lwe_key, (size_t)n * 4)))
        return rc;
    hipError_t e = launch_decrypt((int32_t*)db, d_samples, (const int32_t*)dk, n, count, cx.stream);
    if (e != hipSuccess) return hip_fail("decrypt launch", e);
    return d2h_sync(bits, db, (size_t)count * 4);
}

// ---- attainable VALU rate (roofline denominator check) ----
extern "C" int tfhe_gpu_measure_valu_peak(double* fp64_tflops, double* fp32_tflops) {
    Ctx& c = ctx();
    std::lock_guard<std::recursive_mutex> g(c.mu);
    int rc = ensure_init();
    if (rc) return rc;
    int cus = 0;
    HIPR(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device));
    const int blocks = cus * 16, iters = 4096;
    void* buf;
    if ((rc = scratch().get(6, (size_t)blocks * 256 * 8, &buf))) return rc;
    double* outs[2] = {fp32_tflops, fp64_tflops};
    for (int fp64 = 0; fp64 < 2; ++fp64) {
        hipEvent_t e0, e1;
        HIPR(hipEventCreate(&e0));
        HIPR(hipEventCreate(&e1));
        hipError_t e = launch_valu_probe(fp64, buf, blocks, iters, c.stream);  // warm-up
        if (e != hipSuccess) return hip_fail("valu probe", e);
        HIPR(hipEventRecord(e0, c.stream));
        for (int r = 0; r < 3; ++r)
            if ((e = launch_valu_probe(fp64, buf, blocks, iters, c.stream)) != hipSuccess) return hip_fail("valu probe", e);
        HIPR(hipEventRecord(e1, c.stream));
        HIPR(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPR(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        const double flop = 3.0 * blocks * 256.0 * iters * 8 * 2;
        if (outs[fp64]) *outs[fp64] = flop / (ms * 1e-3) / 1e12;
    }
    return TFHE_OK;
}

// ---- profiling ----
extern "C" int tfhe_gpu_profile_enable(int32_t on) {
    Ctx& c = ctx();
    std::lock_guard<std::recursive_mutex> g(c.mu);
    c.prof = on != 0;
    return TFHE_OK;
}
extern "C" int tfhe_gpu_profile_reset(void) {
    Ctx& c = ctx();
    std::lock_guard<std::recursive_mutex> g(c.mu);
    drain_profile(c);
    c.prof_acc.clear();
    return TFHE_OK;
}
extern "C" int tfhe_gpu_profile_read(const char* kernel, double* total_ms, int64_t* launches) {
    Ctx& c = ctx();
    std::lock_guard<std::recursive_mutex> g(c.mu);
    drain_profile(c);
    auto it = c.prof_acc.find(kernel ? kernel : "");
    if (total_ms) *total_ms = it == c.prof_acc.end() ? 0.0 : it->second.first;
    if (launches) *launches = it == c.prof_acc.end() ? 0 : it->second.second;
    return TFHE_OK;
}
