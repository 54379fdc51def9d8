// tfhe_io.cpp -- key-set, parameter-set and ciphertext files behind the upstream tfhe_io names.
//
// The reference's client/server demo moves keys and ciphertexts between programs through these
// calls (test_bench/client/alice.c:19-27,49-54; test_bench/server/cloud.c:42-44,54-57,65-67;
// test_bench/client/verify.c:8-10,19-22).  Upstream's on-disk byte format is not available here,
// so files are written in this library's own format; a file written here is read back here.
//
// Format (little-endian).  Every object is a 4-byte tag followed by a uint32 version (1):
//   "TMPS"  parameter set: int32 n, N, k, l, Bgbit, ks_t, ks_basebit;
//           double lwe alpha_min, lwe alpha_max, tlwe alpha_min, tlwe alpha_max
//   "TMCK"  cloud key set: a parameter set object, then the time-domain bootstrapping key
//           int32 [n][(k+1)l][k+1][N] (row r of TGSW sample i, polynomial q), then the keyswitch
//           key int32 [kN][t][2^basebit][n+1] (mask then b of sample ks[i][j][h])
//   "TMSK"  secret key set: a parameter set object, int32 lwe_key[n], int32 tlwe_key[k][N], then
//           a cloud key set object (which repeats the parameter set)
//   "TMLW"  ciphertext: int32 n, int32 a[n], int32 b, double current_variance
// The FFT-domain key is never stored: it is rebuilt on the device from the time-domain key.
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <mutex>
#include <vector>

#include "internal.h"

using namespace tfhe;

namespace {

constexpr uint32_t kVersion = 1;
constexpr int kMaxKsBasebit = 8;
constexpr uint64_t kMaxKeyElems = 1ull << 30;

std::mutex& reg_mutex() {
    static std::mutex m;
    return m;
}
std::map<const TFheGateBootstrappingCloudKeySet*, const TFheGateBootstrappingParameterSet*>& standalone() {
    static std::map<const TFheGateBootstrappingCloudKeySet*, const TFheGateBootstrappingParameterSet*> m;
    return m;
}
std::map<const TFheGateBootstrappingSecretKeySet*, const TFheGateBootstrappingParameterSet*>& owned() {
    static std::map<const TFheGateBootstrappingSecretKeySet*, const TFheGateBootstrappingParameterSet*> m;
    return m;
}

struct Writer {
    FILE* f;
    bool ok = true;
    void raw(const void* p, size_t bytes) {
        if (ok && bytes && fwrite(p, 1, bytes, f) != bytes) ok = false;
    }
    void tag(const char* t) {
        raw(t, 4);
        raw(&kVersion, 4);
    }
    void i32(int32_t v) { raw(&v, 4); }
    void f64(double v) { raw(&v, 8); }
};

struct Reader {
    FILE* f;
    bool ok = true;
    void raw(void* p, size_t bytes) {
        if (ok && bytes && fread(p, 1, bytes, f) != bytes) ok = false;
    }
    bool tag(const char* t) {
        char got[4];
        uint32_t ver = 0;
        raw(got, 4);
        raw(&ver, 4);
        if (ok && (memcmp(got, t, 4) != 0 || ver != kVersion)) ok = false;
        return ok;
    }
    int32_t i32() {
        int32_t v = 0;
        raw(&v, 4);
        return v;
    }
    double f64() {
        double v = 0;
        raw(&v, 8);
        return v;
    }
};

void write_params(Writer& w, const TFheGateBootstrappingParameterSet* p) {
    tfhe_flat_params f;
    flat_params_of(p, &f);
    w.tag("TMPS");
    for (int32_t v : {f.n, f.N, f.k, f.l, f.Bgbit, f.ks_t, f.ks_basebit}) w.i32(v);
    const TLweParams* t = p->tgsw_params->tlwe_params;
    for (double v : {p->in_out_params->alpha_min, p->in_out_params->alpha_max, t->alpha_min, t->alpha_max}) w.f64(v);
}

TFheGateBootstrappingParameterSet* read_params(Reader& r) {
    if (!r.tag("TMPS")) return nullptr;
    tfhe_flat_params f;
    f.n = r.i32();
    f.N = r.i32();
    f.k = r.i32();
    f.l = r.i32();
    f.Bgbit = r.i32();
    f.ks_t = r.i32();
    f.ks_basebit = r.i32();
    const double a0 = r.f64(), a1 = r.f64(), a2 = r.f64(), a3 = r.f64();
    if (!r.ok) return nullptr;
    // sizes a file may describe: anything the host containers can hold sensibly.  The keyswitch
    // base 2^basebit is bounded before it is formed, and so are the element counts of the two keys
    // the file would then bring (kMaxKeyElems int32 each, 4 GiB), before anything is allocated.
    if (f.n <= 0 || f.n > 4096 || f.N <= 0 || f.N > 65536 || (f.N & (f.N - 1)) || f.k < 1 || f.k > 4 || f.l < 1 ||
        f.l > 16 || f.Bgbit < 1 || f.l * f.Bgbit > 32 || f.ks_t < 1 || f.ks_basebit < 1 ||
        f.ks_basebit > kMaxKsBasebit || f.ks_t * f.ks_basebit > 32) {
        r.ok = false;
        return nullptr;
    }
    const uint64_t bk_elems = (uint64_t)f.n * (f.k + 1) * f.l * (f.k + 1) * f.N;
    const uint64_t ksk_elems = (uint64_t)f.k * f.N * f.ks_t * (1u << f.ks_basebit) * (f.n + 1);
    if (bk_elems > kMaxKeyElems || ksk_elems > kMaxKeyElems) {
        r.ok = false;
        return nullptr;
    }
    return alloc_params(f, a0, a1, a2, a3);
}

void write_cloud(Writer& w, const TFheGateBootstrappingCloudKeySet* c) {
    w.tag("TMCK");
    write_params(w, c->params);
    tfhe_flat_params f;
    flat_params_of(c->params, &f);
    const int kpl = (f.k + 1) * f.l;
    std::vector<int32_t> bk((size_t)f.n * kpl * (f.k + 1) * f.N);
    flatten_bk(c->bk, bk.data());
    w.raw(bk.data(), bk.size() * 4);
    std::vector<int32_t> ks((size_t)f.k * f.N * f.ks_t * (1 << f.ks_basebit) * (f.n + 1));
    flatten_ksk(c->bk->ks, ks.data());
    w.raw(ks.data(), ks.size() * 4);
}

// fills the containers of an allocated cloud key set from the file
bool read_cloud_body(Reader& r, const TFheGateBootstrappingCloudKeySet& c) {
    tfhe_flat_params f;
    flat_params_of(c.params, &f);
    const int kpl = (f.k + 1) * f.l;
    LweBootstrappingKey* bk = const_cast<LweBootstrappingKey*>(c.bk);
    const double bk_var = c.params->tgsw_params->tlwe_params->alpha_min * c.params->tgsw_params->tlwe_params->alpha_min;
    for (int i = 0; i < f.n && r.ok; ++i)
        for (int row = 0; row < kpl && r.ok; ++row) {
            TLweSample* s = bk->bk[i].all_sample + row;
            for (int q = 0; q <= f.k; ++q) r.raw(s->a[q].coefsT, (size_t)f.N * 4);
            s->current_variance = bk_var;
        }
    LweKeySwitchKey* ks = bk->ks;
    const double ks_var = c.params->in_out_params->alpha_min * c.params->in_out_params->alpha_min;
    for (int i = 0; i < ks->n && r.ok; ++i)
        for (int j = 0; j < ks->t && r.ok; ++j)
            for (int h = 0; h < ks->base && r.ok; ++h) {
                LweSample* s = &ks->ks[i][j][h];
                r.raw(s->a, (size_t)f.n * 4);
                r.raw(&s->b, 4);
                s->current_variance = h ? ks_var : 0.;
            }
    return r.ok;
}

bool same_params(const TFheGateBootstrappingParameterSet* a, const TFheGateBootstrappingParameterSet* b) {
    tfhe_flat_params x, y;
    flat_params_of(a, &x);
    flat_params_of(b, &y);
    return memcmp(&x, &y, sizeof x) == 0;
}

void fail_io(const char* what) {
    set_error(what);
    fprintf(stderr, "tfhe_mi355x: %s\n", what);
}

}  // namespace

namespace tfhe {

void note_standalone_cloud(const TFheGateBootstrappingCloudKeySet* c, const TFheGateBootstrappingParameterSet* p) {
    std::lock_guard<std::mutex> g(reg_mutex());
    standalone()[c] = p;
}
bool take_standalone_cloud(const TFheGateBootstrappingCloudKeySet* c, const TFheGateBootstrappingParameterSet** p) {
    std::lock_guard<std::mutex> g(reg_mutex());
    auto it = standalone().find(c);
    if (it == standalone().end()) return false;
    *p = it->second;
    standalone().erase(it);
    return true;
}
void note_owned_params(const TFheGateBootstrappingSecretKeySet* k, const TFheGateBootstrappingParameterSet* p) {
    std::lock_guard<std::mutex> g(reg_mutex());
    owned()[k] = p;
}
const TFheGateBootstrappingParameterSet* take_owned_params(const TFheGateBootstrappingSecretKeySet* k) {
    std::lock_guard<std::mutex> g(reg_mutex());
    auto it = owned().find(k);
    if (it == owned().end()) return nullptr;
    const TFheGateBootstrappingParameterSet* p = it->second;
    owned().erase(it);
    return p;
}

}  // namespace tfhe

// =============================================================================================
// upstream tfhe_io entry points
// =============================================================================================
extern "C" void export_tfheGateBootstrappingParameterSet_toFile(FILE* F, const TFheGateBootstrappingParameterSet* params) {
    Writer w{F};
    if (!F || !params) fatal("export_tfheGateBootstrappingParameterSet_toFile (null argument)", TFHE_E_ARG);
    write_params(w, params);
    if (!w.ok) fatal("export_tfheGateBootstrappingParameterSet_toFile (write)", TFHE_E_ARG);
}

static TFheGateBootstrappingParameterSet* new_tfheGateBootstrappingParameterSet_fromFile_impl(FILE* F) {
    if (!F) return fail_io("null FILE"), nullptr;
    Reader r{F};
    TFheGateBootstrappingParameterSet* p = read_params(r);
    if (!p) fail_io("not a parameter-set object of this library (tag TMPS)");
    return p;
}

extern "C" void export_tfheGateBootstrappingCloudKeySet_toFile(FILE* F, const TFheGateBootstrappingCloudKeySet* bk) {
    if (!F || !bk || !bk->bk) fatal("export_tfheGateBootstrappingCloudKeySet_toFile (null argument)", TFHE_E_ARG);
    Writer w{F};
    write_cloud(w, bk);
    if (!w.ok) fatal("export_tfheGateBootstrappingCloudKeySet_toFile (write)", TFHE_E_ARG);
}

static TFheGateBootstrappingCloudKeySet* new_tfheGateBootstrappingCloudKeySet_fromFile_impl(FILE* F) {
    if (!F) return fail_io("null FILE"), nullptr;
    Reader r{F};
    if (!r.tag("TMCK")) return fail_io("not a cloud-key object of this library (tag TMCK)"), nullptr;
    TFheGateBootstrappingParameterSet* p = read_params(r);
    if (!p) return fail_io("cloud key: bad parameter set"), nullptr;
    TFheGateBootstrappingCloudKeySet* c = new TFheGateBootstrappingCloudKeySet(alloc_cloud_keyset(p));
    note_standalone_cloud(c, p);
    if (!read_cloud_body(r, *c)) {
        delete_gate_bootstrapping_cloud_keyset(c);
        return fail_io("cloud key: truncated file"), nullptr;
    }
    register_keyset(c);
    return c;
}

extern "C" void export_tfheGateBootstrappingSecretKeySet_toFile(FILE* F, const TFheGateBootstrappingSecretKeySet* key) {
    if (!F || !key) fatal("export_tfheGateBootstrappingSecretKeySet_toFile (null argument)", TFHE_E_ARG);
    Writer w{F};
    w.tag("TMSK");
    write_params(w, key->params);
    const int n = key->params->in_out_params->n;
    const TLweParams* t = key->params->tgsw_params->tlwe_params;
    w.raw(key->lwe_key->key, (size_t)n * 4);
    for (int q = 0; q < t->k; ++q) w.raw(key->tgsw_key->tlwe_key.key[q].coefs, (size_t)t->N * 4);
    write_cloud(w, &key->cloud);
    if (!w.ok) fatal("export_tfheGateBootstrappingSecretKeySet_toFile (write)", TFHE_E_ARG);
}

static TFheGateBootstrappingSecretKeySet* new_tfheGateBootstrappingSecretKeySet_fromFile_impl(FILE* F) {
    if (!F) return fail_io("null FILE"), nullptr;
    Reader r{F};
    if (!r.tag("TMSK")) return fail_io("not a secret-key object of this library (tag TMSK)"), nullptr;
    TFheGateBootstrappingParameterSet* p = read_params(r);
    if (!p) return fail_io("secret key: bad parameter set"), nullptr;
    LweKey* lk = alloc_lwe_key(p->in_out_params);
    TGswKey* gk = alloc_tgsw_key(p->tgsw_params);
    const TLweParams* t = p->tgsw_params->tlwe_params;
    r.raw(lk->key, (size_t)p->in_out_params->n * 4);
    for (int q = 0; q < t->k; ++q) r.raw(gk->tlwe_key.key[q].coefs, (size_t)t->N * 4);
    TFheGateBootstrappingSecretKeySet* ks =
        new TFheGateBootstrappingSecretKeySet{p, lk, gk, alloc_cloud_keyset(p)};
    note_owned_params(ks, p);
    bool ok = r.ok && r.tag("TMCK");
    TFheGateBootstrappingParameterSet* p2 = ok ? read_params(r) : nullptr;
    ok = p2 && same_params(p, p2);
    if (p2) delete_gate_bootstrapping_parameters(p2);
    ok = ok && read_cloud_body(r, ks->cloud);
    if (!ok) {
        delete_gate_bootstrapping_secret_keyset(ks);
        return fail_io("secret key: truncated or inconsistent file"), nullptr;
    }
    register_keyset(&ks->cloud);
    return ks;
}

extern "C" void export_gate_bootstrapping_ciphertext_toFile(FILE* F, const LweSample* sample,
                                                            const TFheGateBootstrappingParameterSet* params) {
    if (!F || !sample || !params) fatal("export_gate_bootstrapping_ciphertext_toFile (null argument)", TFHE_E_ARG);
    const int n = params->in_out_params->n;
    Writer w{F};
    w.tag("TMLW");
    w.i32(n);
    w.raw(sample->a, (size_t)n * 4);
    w.i32(sample->b);
    w.f64(sample->current_variance);
    if (!w.ok) fatal("export_gate_bootstrapping_ciphertext_toFile (write)", TFHE_E_ARG);
}

extern "C" void import_gate_bootstrapping_ciphertext_fromFile(FILE* F, LweSample* sample,
                                                              const TFheGateBootstrappingParameterSet* params) {
    if (!F || !sample || !params) fatal("import_gate_bootstrapping_ciphertext_fromFile (null argument)", TFHE_E_ARG);
    const int n = params->in_out_params->n;
    Reader r{F};
    if (!r.tag("TMLW") || r.i32() != n) {
        set_error("not a ciphertext of this parameter set (tag TMLW)");
        fatal("import_gate_bootstrapping_ciphertext_fromFile", TFHE_E_ARG);
    }
    r.raw(sample->a, (size_t)n * 4);
    sample->b = r.i32();
    sample->current_variance = r.f64();
    if (!r.ok) {
        set_error("truncated ciphertext file");
        fatal("import_gate_bootstrapping_ciphertext_fromFile", TFHE_E_ARG);
    }
}

extern "C" TFheGateBootstrappingParameterSet* new_tfheGateBootstrappingParameterSet_fromFile(FILE* F) {
    // a size the parameter checks admit can still exceed host memory: report it, do not throw
    // through the C ABI
    try {
        return new_tfheGateBootstrappingParameterSet_fromFile_impl(F);
    } catch (const std::bad_alloc&) {
        fail_io("out of host memory while reading the file");
        return nullptr;
    }
}

extern "C" TFheGateBootstrappingCloudKeySet* new_tfheGateBootstrappingCloudKeySet_fromFile(FILE* F) {
    // a size the parameter checks admit can still exceed host memory: report it, do not throw
    // through the C ABI
    try {
        return new_tfheGateBootstrappingCloudKeySet_fromFile_impl(F);
    } catch (const std::bad_alloc&) {
        fail_io("out of host memory while reading the file");
        return nullptr;
    }
}

extern "C" TFheGateBootstrappingSecretKeySet* new_tfheGateBootstrappingSecretKeySet_fromFile(FILE* F) {
    // a size the parameter checks admit can still exceed host memory: report it, do not throw
    // through the C ABI
    try {
        return new_tfheGateBootstrappingSecretKeySet_fromFile_impl(F);
    } catch (const std::bad_alloc&) {
        fail_io("out of host memory while reading the file");
        return nullptr;
    }
}
