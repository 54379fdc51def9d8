// circuit.cpp -- boolean circuits evaluated level by level on the device.
//
// The reference evaluates a circuit one gate call at a time on the host
// (test_bench/server/cloud.c:8-40: bootsXNOR / bootsMUX / bootsOR in loops).  Here a netlist is
// levelised once; each level then costs one table-driven bootstrap launch (every gate of the
// level, over every instance of the batch), at most two keyswitch launches (MUX gates, whose two
// bootstraps are summed before the keyswitch, and all other gates), and nothing else: NOT and
// COPY never bootstrap -- they are folded into the gate coefficients of their consumers and only
// materialised for circuit outputs.  The launch sequence of a bound circuit can be captured once
// into a HIP graph and replayed, and the FP64 launch form of each level width can be chosen by
// timing both forms.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "../hip/engine.h"
#include "internal.h"

using namespace tfhe;

namespace {

constexpr int32_t kMU = (int32_t)(1u << 29);  // 1/8 of the torus, the boolean encoding

struct Node {
    int32_t op;       // TFHE_GATE_*, TFHE_OP_NOT, TFHE_OP_COPY, or kOpConst / kOpInput
    int32_t a, b, c;  // input wires (-1 when unused)
    int32_t value;    // constants: 0 or 1
};
constexpr int32_t kOpInput = -1;
constexpr int32_t kOpConst = -2;

// a wire as an affine image of a stored wire: value = mul * base + (0, add)
struct Lin {
    int32_t base;
    int32_t mul, add;
};

struct LevelPlan {
    std::vector<int32_t> mux, other;  // bootstrapped gate wires of the level
    size_t seg_off = 0, ks_mux_off = 0, ks_oth_off = 0;  // byte offsets into the table buffer
    int form = 0;                     // FP64 launch form (BootLaunch::form)
};

}  // namespace

struct tfhe_circuit {
    int32_t n_in = 0;
    std::vector<Node> nodes;  // wire w: nodes[w] (inputs first)
    std::vector<int32_t> outputs;
    // ---- plan (rebuilt when the netlist changes) ----
    bool planned = false;
    std::vector<Lin> lin;
    std::vector<int32_t> level;
    std::vector<LevelPlan> levels;  // levels[0] is level 1
    std::vector<int32_t> consts;
    int64_t n_boot = 0, max_level_boot = 0;
    // ---- binding of the last run ----
    // the bound key: its address and creation number (the address alone can be reused by a key
    // created after this one is freed, and a captured graph holds the old key's device pointers)
    tfhe_gpu_key* key = nullptr;
    uint64_t key_id = 0;
    int32_t prec = -1, batch = 0, flags = 0;
    const int32_t* d_in = nullptr;
    int32_t* d_out = nullptr;
    bool bound = false;
    int32_t* d_wires = nullptr;
    size_t wires_bytes = 0;
    int32_t* d_u = nullptr;
    size_t u_bytes = 0;
    char* d_tab = nullptr;
    size_t tab_bytes = 0;
    size_t const_off = 0, out_off = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
};

namespace {

bool is_boot(int32_t op) { return op >= TFHE_GATE_NAND && op <= TFHE_GATE_MUX; }

int plan(tfhe_circuit* c) {
    if (c->planned) return TFHE_OK;
    const int32_t W = (int32_t)c->nodes.size();
    c->lin.assign(W, Lin{0, 1, 0});
    c->level.assign(W, 0);
    c->consts.clear();
    int max_level = 0;
    for (int32_t w = 0; w < W; ++w) {
        const Node& nd = c->nodes[w];
        if (nd.op == kOpInput || is_boot(nd.op)) {
            c->lin[w] = Lin{w, 1, 0};
        } else if (nd.op == kOpConst) {
            c->lin[w] = Lin{w, 1, 0};
            c->consts.push_back(w);
        } else {  // NOT / COPY: compose with the input's affine map (upstream bootsNOT negates a and b)
            const Lin& s = c->lin[nd.a];
            const int32_t m = (nd.op == TFHE_OP_NOT) ? -1 : 1;
            c->lin[w] = Lin{s.base, (int32_t)((uint32_t)m * (uint32_t)s.mul), (int32_t)((uint32_t)m * (uint32_t)s.add)};
        }
        int lv = 0;
        if (is_boot(nd.op)) {
            for (int32_t in : {nd.a, nd.b, nd.c})
                if (in >= 0) lv = std::max(lv, c->level[c->lin[in].base] + 1);
        } else if (nd.op == TFHE_OP_NOT || nd.op == TFHE_OP_COPY) {
            lv = c->level[c->lin[w].base];
        }
        c->level[w] = lv;
        max_level = std::max(max_level, lv);
    }
    c->levels.assign(max_level, LevelPlan());
    c->n_boot = c->max_level_boot = 0;
    for (int32_t w = 0; w < W; ++w) {
        const Node& nd = c->nodes[w];
        if (!is_boot(nd.op)) continue;
        LevelPlan& L = c->levels[c->level[w] - 1];
        (nd.op == TFHE_GATE_MUX ? L.mux : L.other).push_back(w);
    }
    for (LevelPlan& L : c->levels) {
        const int64_t nb = 2 * (int64_t)L.mux.size() + (int64_t)L.other.size();
        c->n_boot += nb;
        c->max_level_boot = std::max(c->max_level_boot, nb);
    }
    c->planned = true;
    return TFHE_OK;
}

void drop_graph(tfhe_circuit* c) {
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
}

void release_binding(tfhe_circuit* c) {
    drop_graph(c);
    c->bound = false;
}

int grow(void** p, size_t* have, size_t need) {
    if (*have >= need) return TFHE_OK;
    if (*p) {
        hipError_t e = hipFree(*p);
        if (e != hipSuccess) return lib_hip_fail("hipFree", e);
    }
    *p = nullptr;
    *have = 0;
    hipError_t e = hipMalloc(p, need);
    if (e != hipSuccess) return lib_hip_fail("hipMalloc circuit buffer", e);
    *have = need;
    return TFHE_OK;
}

// Device storage and descriptor tables for one (key, precision, batch, in, out) binding.
int bind(tfhe_circuit* c, tfhe_gpu_key* key, int32_t prec, const int32_t* d_in, int32_t* d_out, int32_t batch) {
    const DevKeyView kv = key_view(key);
    const int n = kv.p.n, N = kv.p.N;
    const size_t samp = (size_t)(n + 1);
    const size_t wire_elems = (size_t)batch * samp;
    const int32_t W = (int32_t)c->nodes.size();
    int rc;
    if ((rc = grow((void**)&c->d_wires, &c->wires_bytes, std::max<size_t>(4, (size_t)(W - c->n_in) * wire_elems * 4))))
        return rc;
    if ((rc = grow((void**)&c->d_u, &c->u_bytes, std::max<size_t>(4, (size_t)c->max_level_boot * batch * (N + 1) * 4))))
        return rc;
    auto wire_ptr = [&](int32_t w) -> int32_t* {
        if (w < c->n_in) return const_cast<int32_t*>(d_in) + (size_t)w * wire_elems;
        return c->d_wires + (size_t)(w - c->n_in) * wire_elems;
    };
    // host image of every table, uploaded at once
    std::vector<char> host;
    auto put = [&](const void* p, size_t bytes) {
        size_t off = (host.size() + 15) & ~(size_t)15;
        host.resize(off + bytes);
        if (bytes) memcpy(host.data() + off, p, bytes);
        return off;
    };
    // gate pre-combination x = (0, C) + alpha a + beta b with both inputs taken through their
    // affine maps; a constant-only input never occurs since constants are stored wires
    auto seg = [&](int gate, int32_t wa, int32_t wb) {
        int32_t C, al, be;
        gate_coefficients(gate, &C, &al, &be);
        const Lin& A = c->lin[wa];
        const Lin& B = c->lin[wb];
        SegDesc d{};
        d.a = wire_ptr(A.base);
        d.b = wire_ptr(B.base);
        d.alpha = (int32_t)((uint32_t)al * (uint32_t)A.mul);
        d.beta = (int32_t)((uint32_t)be * (uint32_t)B.mul);
        d.cst = (int32_t)((uint32_t)C + (uint32_t)al * (uint32_t)A.add + (uint32_t)be * (uint32_t)B.add);
        d.pad = 0;
        return d;
    };
    for (LevelPlan& L : c->levels) {
        std::vector<SegDesc> segs;
        std::vector<int32_t*> ks_mux, ks_oth;
        // upstream bootsMUX: AND(a, b) and AND(NOT a, c) bootstrapped without keyswitch, summed with
        // (0, 1/8), then one keyswitch
        for (int32_t w : L.mux) segs.push_back(seg(TFHE_GATE_AND, c->nodes[w].a, c->nodes[w].b));
        for (int32_t w : L.mux) segs.push_back(seg(TFHE_GATE_ANDNY, c->nodes[w].a, c->nodes[w].c));
        for (int32_t w : L.other) segs.push_back(seg(c->nodes[w].op, c->nodes[w].a, c->nodes[w].b));
        for (int32_t w : L.mux) ks_mux.push_back(wire_ptr(w));
        for (int32_t w : L.other) ks_oth.push_back(wire_ptr(w));
        L.seg_off = put(segs.data(), segs.size() * sizeof(SegDesc));
        L.ks_mux_off = put(ks_mux.data(), ks_mux.size() * sizeof(int32_t*));
        L.ks_oth_off = put(ks_oth.data(), ks_oth.size() * sizeof(int32_t*));
    }
    std::vector<LinDesc> cd;
    for (int32_t w : c->consts) cd.push_back(LinDesc{nullptr, wire_ptr(w), 0, c->nodes[w].value ? kMU : -kMU});
    c->const_off = put(cd.data(), cd.size() * sizeof(LinDesc));
    std::vector<LinDesc> od;
    for (size_t i = 0; i < c->outputs.size(); ++i) {
        const Lin& l = c->lin[c->outputs[i]];
        od.push_back(LinDesc{wire_ptr(l.base), d_out + i * wire_elems, l.mul, l.add});
    }
    c->out_off = put(od.data(), od.size() * sizeof(LinDesc));
    if ((rc = grow((void**)&c->d_tab, &c->tab_bytes, std::max<size_t>(16, host.size())))) return rc;
    hipError_t e = hipMemcpyAsync(c->d_tab, host.data(), host.size(), hipMemcpyHostToDevice, lib_stream());
    if (e != hipSuccess) return lib_hip_fail("upload circuit tables", e);
    if ((e = hipStreamSynchronize(lib_stream())) != hipSuccess) return lib_hip_fail("circuit tables sync", e);
    release_binding(c);
    c->key = key;
    c->key_id = key_id(key);
    c->prec = prec;
    c->batch = batch;
    c->d_in = d_in;
    c->d_out = d_out;
    c->bound = true;
    return TFHE_OK;
}

// launch form per (N, precision, bootstraps in the level) chosen by timing, shared by every
// circuit of the process
std::map<std::tuple<int, int, int64_t>, int>& tuned_forms() {
    static std::map<std::tuple<int, int, int64_t>, int> m;
    return m;
}

int enqueue_level_boot(tfhe_circuit* c, const LevelPlan& L, DevTwiddles* tw, int form) {
    const DevKeyView kv = key_view(c->key);
    const int jobs_per_seg = c->batch;
    const int nseg = (int)(2 * L.mux.size() + L.other.size());
    BootLaunch B;
    B.mode = 3;
    B.form = form;
    B.segs = (const SegDesc*)(c->d_tab + L.seg_off);
    B.seg_jobs = jobs_per_seg;
    B.count = nseg * jobs_per_seg;
    B.mu = kMU;
    B.u_out = c->d_u;
    LibProf pr(c->prec == TFHE_PREC_FP64 ? "bootstrap_fp64" : "bootstrap_split32");
    hipError_t e = launch_bootstrap(kv, *tw, (Prec)c->prec, B, lib_stream());
    return e == hipSuccess ? TFHE_OK : lib_hip_fail("circuit bootstrap launch", e);
}

int enqueue_all(tfhe_circuit* c, DevTwiddles* tw) {
    const DevKeyView kv = key_view(c->key);
    const int n = kv.p.n, N = kv.p.N;
    const int B = c->batch;
    hipStream_t st = lib_stream();
    hipError_t e;
    if (!c->consts.empty() &&
        (e = launch_lwe_linear((const LinDesc*)(c->d_tab + c->const_off), (int)c->consts.size(), B, n, st)) != hipSuccess)
        return lib_hip_fail("circuit constants", e);
    for (const LevelPlan& L : c->levels) {
        if (L.mux.empty() && L.other.empty()) continue;
        int rc = enqueue_level_boot(c, L, tw, L.form);
        if (rc) return rc;
        LibProf pr("keyswitch");
        const size_t m = L.mux.size() * (size_t)B;
        if (m) {
            e = launch_keyswitch(kv, nullptr, c->d_u, c->d_u + m * (N + 1), kMU, (int)m, st,
                                 (int32_t* const*)(c->d_tab + L.ks_mux_off), B);
            if (e != hipSuccess) return lib_hip_fail("circuit keyswitch (MUX)", e);
        }
        if (!L.other.empty()) {
            e = launch_keyswitch(kv, nullptr, c->d_u + 2 * m * (N + 1), nullptr, 0, (int)(L.other.size() * B), st,
                                 (int32_t* const*)(c->d_tab + L.ks_oth_off), B);
            if (e != hipSuccess) return lib_hip_fail("circuit keyswitch", e);
        }
    }
    if (!c->outputs.empty() &&
        (e = launch_lwe_linear((const LinDesc*)(c->d_tab + c->out_off), (int)c->outputs.size(), B, n, st)) != hipSuccess)
        return lib_hip_fail("circuit outputs", e);
    return TFHE_OK;
}

// Launch forms worth timing for a level of `jobs` bootstraps (BootLaunch::form):
//   N = 1024, FP64:  1 key-sharing workgroups, 2 one bootstrap per workgroup;
//   N = 1024, split: 2 one bootstrap per workgroup (its only batched form);
//   N = 2048:        2, 4, 5 the two-wavefront kernel with 1, 2, 4 bootstraps per workgroup;
//   and, for levels of at most kLatencyMaxJobs bootstraps, 3 the latency form (one bootstrap per
//   workgroup of several wavefronts).
constexpr int64_t kLatencyMaxJobs = 1024;
constexpr int kTuneSamples = 3;

std::vector<int> tune_candidates(int N, int32_t prec, int64_t jobs) {
    std::vector<int> c;
    if (N == 1024) c = prec == TFHE_PREC_FP64 ? std::vector<int>{1, 2} : std::vector<int>{2};
    else c = {2, 4, 5};
    if (jobs <= kLatencyMaxJobs) c.push_back(3);
    return c;
}

// Times the candidate launch forms of each level width not seen before: per form one untimed
// launch (warm-up), then kTuneSamples launches each bracketed by events; the form with the lowest
// median is kept.  The levels run on whatever the wire storage holds (only the timing is used);
// the real evaluation follows.
int autotune(tfhe_circuit* c, DevTwiddles* tw) {
    const DevKeyView kv = key_view(c->key);
    hipStream_t st = lib_stream();
    for (LevelPlan& L : c->levels) {
        L.form = 0;
        const int64_t jobs = (int64_t)(2 * L.mux.size() + L.other.size()) * c->batch;
        if (!jobs) continue;
        const std::vector<int> cand = tune_candidates(kv.p.N, c->prec, jobs);
        if (cand.size() < 2) {
            L.form = cand.empty() ? 0 : cand[0];
            continue;
        }
        auto key = std::make_tuple(kv.p.N, (int)c->prec, jobs);
        auto it = tuned_forms().find(key);
        if (it == tuned_forms().end()) {
            hipEvent_t ev[kTuneSamples + 1];
            for (hipEvent_t& e : ev)
                if (hipEventCreate(&e) != hipSuccess) return lib_hip_fail("event", hipErrorUnknown);
            float best = 0.f;
            int best_form = cand[0];
            int rc = TFHE_OK;
            for (int form : cand) {
                if ((rc = enqueue_level_boot(c, L, tw, form))) break;   // warm-up
                for (int r = 0; r < kTuneSamples && !rc; ++r) {
                    (void)hipEventRecord(ev[r], st);
                    rc = enqueue_level_boot(c, L, tw, form);
                }
                (void)hipEventRecord(ev[kTuneSamples], st);
                if (rc) break;
                hipError_t e = hipEventSynchronize(ev[kTuneSamples]);
                if (e != hipSuccess) {
                    rc = lib_hip_fail("autotune sync", e);
                    break;
                }
                float ms[kTuneSamples];
                for (int r = 0; r < kTuneSamples; ++r) (void)hipEventElapsedTime(&ms[r], ev[r], ev[r + 1]);
                std::sort(ms, ms + kTuneSamples);
                const float med = ms[kTuneSamples / 2];
                if (form == cand[0] || med < best) {
                    best = med;
                    best_form = form;
                }
            }
            for (hipEvent_t& e : ev) (void)hipEventDestroy(e);
            if (rc) return rc;
            it = tuned_forms().emplace(key, best_form).first;
        }
        L.form = it->second;
    }
    return TFHE_OK;
}

}  // namespace

// =============================================================================================
// C ABI
// =============================================================================================
extern "C" int tfhe_circuit_create(int32_t n_inputs, tfhe_circuit** out) {
    if (!out) return lib_arg_fail("null output pointer");
    if (n_inputs < 0) return lib_arg_fail("negative input count");
    tfhe_circuit* c = new tfhe_circuit();
    c->n_in = n_inputs;
    c->nodes.assign(n_inputs, Node{kOpInput, -1, -1, -1, 0});
    *out = c;
    return TFHE_OK;
}

extern "C" int tfhe_circuit_free(tfhe_circuit* c) {
    if (!c) return TFHE_OK;
    std::lock_guard<std::recursive_mutex> g(lib_mutex());
    release_binding(c);
    if (c->d_wires) (void)hipFree(c->d_wires);
    if (c->d_u) (void)hipFree(c->d_u);
    if (c->d_tab) (void)hipFree(c->d_tab);
    delete c;
    return TFHE_OK;
}

extern "C" int tfhe_circuit_add(tfhe_circuit* c, int32_t op, int32_t a, int32_t b, int32_t cin, int32_t* out_wire) {
    if (!c || !out_wire) return lib_arg_fail("null argument");
    const int32_t W = (int32_t)c->nodes.size();
    auto ok = [&](int32_t w) { return w >= 0 && w < W; };
    if (is_boot(op)) {
        if (!ok(a) || !ok(b)) return lib_arg_fail("gate input wire out of range");
        if (op == TFHE_GATE_MUX && !ok(cin)) return lib_arg_fail("MUX third input wire out of range");
        if (op != TFHE_GATE_MUX) cin = -1;
    } else if (op == TFHE_OP_NOT || op == TFHE_OP_COPY) {
        if (!ok(a)) return lib_arg_fail("input wire out of range");
        b = cin = -1;
    } else {
        return lib_arg_fail("unknown circuit op");
    }
    c->nodes.push_back(Node{op, a, b, cin, 0});
    c->planned = false;
    c->bound = false;
    *out_wire = W;
    return TFHE_OK;
}

extern "C" int tfhe_circuit_constant(tfhe_circuit* c, int32_t value, int32_t* out_wire) {
    if (!c || !out_wire) return lib_arg_fail("null argument");
    if (value != 0 && value != 1) return lib_arg_fail("constant must be 0 or 1");
    *out_wire = (int32_t)c->nodes.size();
    c->nodes.push_back(Node{kOpConst, -1, -1, -1, value});
    c->planned = false;
    c->bound = false;
    return TFHE_OK;
}

extern "C" int tfhe_circuit_output(tfhe_circuit* c, int32_t wire) {
    if (!c) return lib_arg_fail("null circuit");
    if (wire < 0 || wire >= (int32_t)c->nodes.size()) return lib_arg_fail("output wire out of range");
    c->outputs.push_back(wire);
    c->bound = false;
    return TFHE_OK;
}

extern "C" int tfhe_circuit_info(tfhe_circuit* c, int32_t* n_levels, int32_t* max_level_bootstraps,
                                 int64_t* n_bootstraps, int32_t* n_outputs) {
    if (!c) return lib_arg_fail("null circuit");
    plan(c);
    if (n_levels) *n_levels = (int32_t)c->levels.size();
    if (max_level_bootstraps) *max_level_bootstraps = (int32_t)c->max_level_boot;
    if (n_bootstraps) *n_bootstraps = c->n_boot;
    if (n_outputs) *n_outputs = (int32_t)c->outputs.size();
    return TFHE_OK;
}

extern "C" int tfhe_circuit_level_of(tfhe_circuit* c, int32_t wire, int32_t* level) {
    if (!c || !level) return lib_arg_fail("null argument");
    if (wire < 0 || wire >= (int32_t)c->nodes.size()) return lib_arg_fail("wire out of range");
    plan(c);
    *level = c->level[wire];
    return TFHE_OK;
}

extern "C" int tfhe_circuit_level_plan(tfhe_circuit* c, int32_t level, int32_t* bootstraps, int32_t* form) {
    if (!c) return lib_arg_fail("null circuit");
    plan(c);
    if (level < 1 || level > (int32_t)c->levels.size()) return lib_arg_fail("level out of range");
    const LevelPlan& L = c->levels[level - 1];
    if (bootstraps) *bootstraps = (int32_t)(2 * L.mux.size() + L.other.size());
    if (form) *form = L.form;
    return TFHE_OK;
}

extern "C" int tfhe_circuit_run(tfhe_circuit* c, tfhe_gpu_key* key, int32_t prec, const int32_t* d_in,
                                int32_t* d_out, int32_t batch, int32_t flags) {
    std::lock_guard<std::recursive_mutex> g(lib_mutex());
    if (!c) return lib_arg_fail("null circuit");
    if (batch < 0) return lib_arg_fail("negative batch");
    if (batch == 0) return TFHE_OK;
    if ((c->n_in && !d_in) || (!c->outputs.empty() && !d_out)) return lib_arg_fail("null ciphertext array");
    if (flags & ~(TFHE_CIRCUIT_GRAPH | TFHE_CIRCUIT_AUTOTUNE)) return lib_arg_fail("unknown circuit flags");
    int rc = lib_ensure_init();
    if (rc) return rc;
    if ((rc = key_check_prec(key, prec))) return rc;
    const DevKeyView kv = key_view(key);
    DevTwiddles* tw;
    if ((rc = lib_twiddles(kv.p.N, &tw))) return rc;
    if ((rc = plan(c))) return rc;
    if ((int64_t)c->max_level_boot * batch > (int64_t)INT32_MAX / 2) return lib_arg_fail("level too wide for one launch");
    const bool rebind = !c->bound || c->key != key || c->key_id != key_id(key) || c->prec != prec || c->batch != batch || c->d_in != d_in ||
                        c->d_out != d_out;
    if (rebind && (rc = bind(c, key, prec, d_in, d_out, batch))) return rc;
    const bool tune = (flags & TFHE_CIRCUIT_AUTOTUNE) != 0;
    if (rebind || tune != ((c->flags & TFHE_CIRCUIT_AUTOTUNE) != 0)) {
        // launch forms changed: a captured graph holds the old ones
        drop_graph(c);
        if (tune) {
            if ((rc = autotune(c, tw))) return rc;
        } else {
            for (LevelPlan& L : c->levels) L.form = 0;
        }
    }
    c->flags = flags;
    hipStream_t st = lib_stream();
    hipError_t e;
    if (!(flags & TFHE_CIRCUIT_GRAPH)) return enqueue_all(c, tw);
    if (!c->gexec) {
        if ((e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal)) != hipSuccess)
            return lib_hip_fail("begin capture", e);
        rc = enqueue_all(c, tw);
        hipGraph_t gr = nullptr;
        e = hipStreamEndCapture(st, &gr);
        if (rc) {
            if (gr) (void)hipGraphDestroy(gr);
            return rc;
        }
        if (e != hipSuccess) return lib_hip_fail("end capture", e);
        c->graph = gr;
        if ((e = hipGraphInstantiate(&c->gexec, gr, nullptr, nullptr, 0)) != hipSuccess) {
            c->gexec = nullptr;
            return lib_hip_fail("graph instantiate", e);
        }
    }
    if ((e = hipGraphLaunch(c->gexec, st)) != hipSuccess) return lib_hip_fail("graph launch", e);
    return TFHE_OK;
}

extern "C" int tfhe_gpu_circuit_level(tfhe_gpu_key* key, int32_t prec, int32_t n_gates, const int32_t* ops,
                                      const int32_t* const* a, const int32_t* const* b, const int32_t* const* cc,
                                      int32_t* const* out, int32_t batch) {
    // the gates of one level, each over `batch` samples, in one table-driven bootstrap launch and
    // at most two keyswitch launches; scratch is allocated per call and the call is synchronous
    std::lock_guard<std::recursive_mutex> g(lib_mutex());
    if (n_gates < 0 || batch < 0) return lib_arg_fail("negative size");
    if (n_gates == 0 || batch == 0) return TFHE_OK;
    if (!ops || !a || !b || !out) return lib_arg_fail("null argument");
    int rc = lib_ensure_init();
    if (rc) return rc;
    if ((rc = key_check_prec(key, prec))) return rc;
    for (int i = 0; i < n_gates; ++i) {
        if (!is_boot(ops[i])) return lib_arg_fail("circuit level: ops must be two-input gates or MUX");
        if (!a[i] || !b[i] || !out[i] || (ops[i] == TFHE_GATE_MUX && (!cc || !cc[i])))
            return lib_arg_fail("circuit level: null ciphertext array");
    }
    const DevKeyView kv = key_view(key);
    DevTwiddles* tw;
    if ((rc = lib_twiddles(kv.p.N, &tw))) return rc;
    const int N = kv.p.N;
    size_t nmux = 0;
    for (int i = 0; i < n_gates; ++i) nmux += ops[i] == TFHE_GATE_MUX;
    const size_t nboot = (2 * nmux + (n_gates - nmux)) * (size_t)batch;
    void* d_u = nullptr;
    void* d_tab = nullptr;
    std::vector<SegDesc> segs;
    std::vector<int32_t*> ks_mux, ks_oth;
    auto seg = [&](int gate, const int32_t* x, const int32_t* y) {
        SegDesc d{};
        gate_coefficients(gate, &d.cst, &d.alpha, &d.beta);
        d.a = x;
        d.b = y;
        return d;
    };
    for (int i = 0; i < n_gates; ++i)
        if (ops[i] == TFHE_GATE_MUX) { segs.push_back(seg(TFHE_GATE_AND, a[i], b[i])); ks_mux.push_back(out[i]); }
    for (int i = 0; i < n_gates; ++i)
        if (ops[i] == TFHE_GATE_MUX) segs.push_back(seg(TFHE_GATE_ANDNY, a[i], cc[i]));
    for (int i = 0; i < n_gates; ++i)
        if (ops[i] != TFHE_GATE_MUX) { segs.push_back(seg(ops[i], a[i], b[i])); ks_oth.push_back(out[i]); }
    const size_t seg_bytes = segs.size() * sizeof(SegDesc);
    const size_t tab_bytes = seg_bytes + (ks_mux.size() + ks_oth.size()) * sizeof(int32_t*);
    hipStream_t st = lib_stream();
    hipError_t e;
    if ((e = hipMalloc(&d_u, nboot * (N + 1) * 4)) != hipSuccess) return lib_hip_fail("hipMalloc level u", e);
    if ((e = hipMalloc(&d_tab, tab_bytes)) != hipSuccess) {
        (void)hipFree(d_u);
        return lib_hip_fail("hipMalloc level tables", e);
    }
    std::vector<char> host(tab_bytes);
    memcpy(host.data(), segs.data(), seg_bytes);
    memcpy(host.data() + seg_bytes, ks_mux.data(), ks_mux.size() * sizeof(int32_t*));
    memcpy(host.data() + seg_bytes + ks_mux.size() * sizeof(int32_t*), ks_oth.data(), ks_oth.size() * sizeof(int32_t*));
    rc = TFHE_OK;
    BootLaunch B;
    B.mode = 3;
    B.segs = (const SegDesc*)d_tab;
    B.seg_jobs = batch;
    B.count = (int)nboot;
    B.mu = kMU;
    B.u_out = (int32_t*)d_u;
    int32_t* u = (int32_t*)d_u;
    const size_t m = nmux * (size_t)batch;
    if ((e = hipMemcpyAsync(d_tab, host.data(), tab_bytes, hipMemcpyHostToDevice, st)) != hipSuccess)
        rc = lib_hip_fail("upload level tables", e);
    if (!rc && (e = launch_bootstrap(kv, *tw, (Prec)prec, B, st)) != hipSuccess)
        rc = lib_hip_fail("level bootstrap", e);
    if (!rc && m &&
        (e = launch_keyswitch(kv, nullptr, u, u + m * (N + 1), kMU, (int)m, st,
                              (int32_t* const*)((char*)d_tab + seg_bytes), batch)) != hipSuccess)
        rc = lib_hip_fail("level keyswitch (MUX)", e);
    if (!rc && !ks_oth.empty() &&
        (e = launch_keyswitch(kv, nullptr, u + 2 * m * (N + 1), nullptr, 0, (int)(ks_oth.size() * batch), st,
                              (int32_t* const*)((char*)d_tab + seg_bytes + ks_mux.size() * sizeof(int32_t*)),
                              batch)) != hipSuccess)
        rc = lib_hip_fail("level keyswitch", e);
    if ((e = hipStreamSynchronize(st)) != hipSuccess && !rc) rc = lib_hip_fail("level sync", e);
    (void)hipFree(d_u);
    (void)hipFree(d_tab);
    return rc;
}
