// Request server: circuits of mixed size submitted from any thread, executed
// on one device with continuous batching.
//
// No reference counterpart: the reference runs one circuit at a time through
// QuantumCircuit::update_quantum_state (src/cppsim/circuit.cpp:46-57), then
// Observable::get_expectation_value and QuantumStateBase::sampling
// (src/cppsim/state.hpp:899-931).  Each request here returns the same three
// results: its final state's observable expectation and its samples for the
// given uniforms.
//
// One scheduler thread owns the device work:
//   * admission: requests are admitted in arrival order while the sum of their
//     state footprints (16 * 2^n bytes) stays within the budget; a request that
//     does not fit lets later, smaller ones pass until it has waited
//     kStarveSeconds, after which admission waits for it;
//   * small requests (n <= small_max_qubits) join the next batch; whenever the
//     batch stream is idle, every admitted small request is compiled into one
//     batched-executor launch (states in LDS) followed by one launch that builds
//     every state's exact CDF and searches its uniforms;
//   * large requests take one of `large_slots` streams: stream-ordered
//     allocation, |0>, the fused stage executor, the observable, sampling, free;
//   * completion is observed through events; the host copies go to pinned
//     buffers, so nothing waits for a stream.
// Results do not depend on what else is in flight: a circuit's batch program and
// reductions depend only on the circuit, and a large request runs alone on its
// stream.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "qm_internal.hpp"

namespace qm {
namespace {

using clk = std::chrono::steady_clock;
constexpr double kStarveSeconds = 0.25;
// register-block width of the batched executor, the fastest measured for
// brickwork / layered circuits; large requests take the fused executor's
// defaults, which qm_circuit_create picks by qubit count
constexpr u32 kBatchBlockQubits = 3;

struct Request {
    long long id = 0;
    u32 n = 0;
    std::vector<qm_batch_gate> gates;
    std::vector<std::vector<qm_complex>> mats;
    std::vector<u64> bf, pf;
    std::vector<u32> rot;
    std::vector<double> coef;
    std::vector<double> uniforms;
    qm_circuit* circuit = nullptr;  // large requests
    bool small = false;
    u64 bytes = 0;
    clk::time_point t_submit, t_admit, t_done;
    double energy = 0.0;
    double margin = 0.0;  // boundary check of the samples (qm_sampling_margin of the gate count)
    std::vector<u64> samples;
    std::vector<unsigned char> flags;
    int status = 0;  // 0 queued, 1 admitted, 2 done, 3 failed
    std::string error;
    ~Request() {
        if (circuit) qm_circuit_destroy(circuit);
    }
};
using ReqPtr = std::shared_ptr<Request>;

struct PinnedBuf {
    void* p = nullptr;
    size_t cap = 0;
    bool reserve(size_t bytes) {
        if (cap >= bytes && p) return true;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        if (!hip_ok(hipHostMalloc(&p, std::max<size_t>(bytes, 64)), "request server: pinned buffer")) return false;
        cap = std::max<size_t>(bytes, 64);
        return true;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
};

struct Slot {
    void* sptr = nullptr;  // from allocate_cuda_stream_host (owns the stream's scratch)
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    ReqPtr req;
    PinnedBuf energy, samples, uniforms, flags;
};

double seconds(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

u64 large_state_bytes(u32 n) { return ((16ULL << n) + 255) & ~255ULL; }

}  // namespace
}  // namespace qm

using namespace qm;

struct qm_server {
    unsigned device = 0;
    u64 budget = 0;
    u32 small_max = 16;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<ReqPtr> inbox;
    std::map<long long, ReqPtr> table;  // every request not yet collected by wait
    long long next_id = 0;
    bool stop = false;
    std::thread worker;

    // scheduler-owned
    std::deque<ReqPtr> queue;
    std::vector<Slot> slots;
    void* bsptr = nullptr;
    hipStream_t bstream = nullptr;
    hipEvent_t bev = nullptr;
    qm_batch* batch = nullptr;
    std::vector<ReqPtr> running_batch, next_batch;
    PinnedBuf benergy, bsamples, buniforms, bflags;
    u64 in_use = 0;

    // statistics (under mu)
    u64 n_admitted = 0, n_done = 0, n_failed = 0, n_batches = 0, max_batch = 0, peak_bytes = 0;

    void run();
    bool busy() const {
        if (!running_batch.empty() || !next_batch.empty() || !queue.empty()) return true;
        for (const Slot& s : slots)
            if (s.req) return true;
        return false;
    }
    void finish(const ReqPtr& r, bool ok, const std::string& err) {
        std::lock_guard<std::mutex> lk(mu);
        r->t_done = clk::now();
        r->status = ok ? 2 : 3;
        r->error = err;
        in_use -= r->bytes;
        if (ok) ++n_done;
        else ++n_failed;
        cv_done.notify_all();
    }
    std::string take_error() {
        char buf[512];
        const int code = qm_get_last_error(buf, sizeof buf);
        qm_clear_last_error();
        return code == QM_OK ? std::string() : std::string(buf);
    }
    void launch_large(Slot& s, const ReqPtr& r);
    void launch_batch();
    void complete_batch(bool ok, const std::string& err);
};

void qm_server::launch_large(Slot& s, const ReqPtr& r) {
    const u64 dim = 1ULL << r->n;
    const u32 count = (u32)r->uniforms.size();
    std::string err;
    // One stream-ordered allocation holds the state and the sampling CDF with
    // its tables: r->bytes, the amount admission counted for this request.
    void* mem = nullptr;
    bool ok = s.energy.reserve(4 * sizeof(double)) && s.samples.reserve((size_t)count * 8 + 8) &&
              s.uniforms.reserve((size_t)count * 8 + 8) && s.flags.reserve((size_t)count + 8);
    if (ok) {
        std::memcpy(s.uniforms.p, r->uniforms.data(), (size_t)count * 8);
        ok = hip_ok(hipMallocAsync(&mem, r->bytes, s.stream), "request server: state memory");
    }
    if (ok) {
        void* state = mem;
        void* samp = (char*)mem + large_state_bytes(r->n);
        enqueue_basis(state, dim, 0, s.stream);
        ok = circuit_enqueue(r->circuit, state, dim, s.stream, device) &&
             enqueue_observable(r->bf.data(), r->pf.data(), r->rot.data(), r->coef.data(), (u32)r->coef.size(), state,
                                dim, s.stream, (double*)s.energy.p) &&
             enqueue_sampling_in(samp, state, dim, (const double*)s.uniforms.p, count, (u64*)s.samples.p, s.stream,
                                 r->margin, (unsigned char*)s.flags.p);
        (void)hipFreeAsync(mem, s.stream);
    }
    if (!ok) err = take_error();
    s.req = r;
    if (!hip_ok(hipEventRecord(s.ev, s.stream), "request server: event")) {
        if (err.empty()) err = take_error();
        (void)hipStreamSynchronize(s.stream);
        s.req.reset();
        finish(r, false, err);
        return;
    }
    if (!err.empty()) r->error = err;  // reported when the event completes
}

void qm_server::launch_batch() {
    std::string err;
    qm_batch_clear(batch);
    size_t total = 0;
    for (const ReqPtr& r : next_batch) {
        if (qm_batch_add_circuit(batch, r->n, r->gates.data(), (UINT)r->gates.size(), r->bf.data(), r->pf.data(),
                                 r->rot.data(), r->coef.data(), (UINT)r->coef.size()) < 0) {
            err = take_error();
            break;
        }
        total += r->uniforms.size();
    }
    const size_t nb = next_batch.size();
    bool ok = err.empty() && qm_batch_compile(batch, kBatchBlockQubits, 1, device) == QM_OK &&
              benergy.reserve(nb * sizeof(double)) && bsamples.reserve(total * 8 + 8) &&
              buniforms.reserve(total * 8 + 8) && bflags.reserve(total + 8);
    const double* d_energy = nullptr;
    if (ok) {
        d_energy = qm_batch_run_async(batch, bsptr, device);
        ok = d_energy != nullptr &&
             hip_ok(hipMemcpyAsync(benergy.p, d_energy, nb * sizeof(double), hipMemcpyDeviceToHost, bstream),
                    "request server: energies");
    }
    if (ok) {
        std::vector<SampleJob> jobs(nb);
        size_t off = 0;
        for (size_t i = 0; i < nb; ++i) {
            const ReqPtr& r = next_batch[i];
            std::memcpy((double*)buniforms.p + off, r->uniforms.data(), r->uniforms.size() * 8);
            jobs[i].state = (const double2*)qm_batch_state_device_ptr(batch, (UINT)i);
            jobs[i].dim = 1ULL << r->n;
            jobs[i].h_uniforms = (const double*)buniforms.p + off;
            jobs[i].h_out = (u64*)bsamples.p + off;
            jobs[i].count = (u32)r->uniforms.size();
            jobs[i].margin = r->margin;
            jobs[i].h_flags = (unsigned char*)bflags.p + off;
            off += r->uniforms.size();
        }
        ok = enqueue_sampling_jobs(jobs.data(), (u32)nb, bstream);
    }
    if (!ok && err.empty()) err = take_error();
    running_batch.swap(next_batch);
    next_batch.clear();
    {
        std::lock_guard<std::mutex> lk(mu);
        ++n_batches;
        max_batch = std::max<u64>(max_batch, running_batch.size());
    }
    if (!err.empty() || !hip_ok(hipEventRecord(bev, bstream), "request server: event")) {
        if (err.empty()) err = take_error();
        (void)hipStreamSynchronize(bstream);
        complete_batch(false, err);
    }
}

void qm_server::complete_batch(bool ok, const std::string& err) {
    size_t off = 0;
    for (size_t i = 0; i < running_batch.size(); ++i) {
        const ReqPtr& r = running_batch[i];
        if (ok) {
            r->energy = ((const double*)benergy.p)[i];
            r->samples.assign((const u64*)bsamples.p + off, (const u64*)bsamples.p + off + r->uniforms.size());
            r->flags.assign((const unsigned char*)bflags.p + off,
                            (const unsigned char*)bflags.p + off + r->uniforms.size());
        }
        off += r->uniforms.size();
        finish(r, ok, err);
    }
    running_batch.clear();
}

void qm_server::run() {
    if (!use_device(device)) return;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(mu);
            if (inbox.empty() && !busy()) {
                if (stop) break;
                cv_work.wait_for(lk, std::chrono::milliseconds(2));
            }
            while (!inbox.empty()) {
                queue.push_back(inbox.front());
                inbox.pop_front();
            }
        }
        bool progress = false;
        // completions
        for (Slot& s : slots) {
            if (!s.req) continue;
            const hipError_t q = hipEventQuery(s.ev);
            if (q == hipErrorNotReady) continue;
            ReqPtr r = s.req;
            s.req.reset();
            const bool ok = q == hipSuccess && r->error.empty();
            if (ok) {
                const double* e = (const double*)s.energy.p;
                r->energy = e[0] + e[2];
                r->samples.assign((const u64*)s.samples.p, (const u64*)s.samples.p + r->uniforms.size());
                r->flags.assign((const unsigned char*)s.flags.p,
                                (const unsigned char*)s.flags.p + r->uniforms.size());
            }
            finish(r, ok, ok ? std::string() : (r->error.empty() ? std::string(hipGetErrorString(q)) : r->error));
            progress = true;
        }
        if (!running_batch.empty()) {
            const hipError_t q = hipEventQuery(bev);
            if (q != hipErrorNotReady) {
                complete_batch(q == hipSuccess, q == hipSuccess ? std::string() : std::string(hipGetErrorString(q)));
                progress = true;
            }
        }
        // admission
        bool head_blocked = false;
        for (auto it = queue.begin(); it != queue.end();) {
            const ReqPtr& r = *it;
            bool admit = false;
            {
                std::lock_guard<std::mutex> lk(mu);
                const bool fits = in_use + r->bytes <= budget;
                if (fits && !r->small) {
                    bool free_slot = false;
                    for (const Slot& s : slots) free_slot = free_slot || !s.req;
                    admit = free_slot;
                } else {
                    admit = fits;
                }
                if (admit) {
                    in_use += r->bytes;
                    peak_bytes = std::max(peak_bytes, in_use);
                    ++n_admitted;
                    r->status = 1;
                    r->t_admit = clk::now();
                }
            }
            if (!admit) {
                if (it == queue.begin() && seconds(r->t_submit, clk::now()) > kStarveSeconds) head_blocked = true;
                if (head_blocked) break;  // let memory drain for the starving head
                ++it;
                continue;
            }
            ReqPtr keep = r;
            it = queue.erase(it);
            progress = true;
            if (keep->small) {
                next_batch.push_back(keep);
            } else {
                for (Slot& s : slots)
                    if (!s.req) {
                        launch_large(s, keep);
                        break;
                    }
            }
        }
        if (running_batch.empty() && !next_batch.empty()) {
            launch_batch();
            progress = true;
        }
        if (!progress) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

extern "C" {

qm_server* qm_server_create(ITYPE hbm_budget_bytes, UINT large_slots, UINT small_max_qubits, unsigned device_number) {
    const char* what = "qm_server_create";
    if (hbm_budget_bytes < 4096 || large_slots < 1 || large_slots > 16 || small_max_qubits > 16) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: need a budget >= 4 KiB, 1..16 large slots, small_max_qubits <= 16", what);
        return nullptr;
    }
    if (!use_device(device_number)) return nullptr;
    auto* s = new qm_server();
    s->device = device_number;
    s->budget = hbm_budget_bytes;
    s->small_max = small_max_qubits;
    s->slots.resize(large_slots);
    s->bsptr = allocate_cuda_stream_host(1, device_number);
    bool ok = s->bsptr != nullptr && hip_ok(hipEventCreateWithFlags(&s->bev, hipEventDisableTiming), what);
    if (ok) s->bstream = stream_of(s->bsptr);
    for (Slot& sl : s->slots) {
        if (!ok) break;
        sl.sptr = allocate_cuda_stream_host(1, device_number);
        ok = sl.sptr != nullptr && hip_ok(hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming), what);
        if (ok) sl.stream = stream_of(sl.sptr);
    }
    s->batch = qm_batch_create();
    if (!ok) {
        qm_server_destroy(s);
        return nullptr;
    }
    s->worker = std::thread([s] { s->run(); });
    return s;
}

long long qm_server_submit(qm_server* s, UINT n, const qm_batch_gate* gates, UINT gate_count, const ITYPE* bf,
                           const ITYPE* pf, const UINT* rot, const double* coefs, UINT term_count,
                           const double* uniforms, UINT sample_count) {
    const char* what = "qm_server_submit";
    if (!s || n < 1 || n > QM_CIRCUIT_MAX_QUBITS || (gate_count && !gates) ||
        (term_count && (!bf || !pf || !rot || !coefs)) || (sample_count && !uniforms)) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: bad server, qubit count or null array", what);
        return -1;
    }
    auto r = std::make_shared<Request>();
    r->n = n;
    r->small = n <= std::max<u32>(s->small_max, 3);  // the fused executor needs >= 4 qubits
    // HBM footprint: a small request's state (16 * 2^n; its CDF lives in the
    // batch stream's scratch, at most 8 * (2^16 + 1) bytes per request); a large
    // request's one allocation of state + CDF + CDF tables + sample buffers.
    r->bytes = r->small ? (16ULL << n) : large_state_bytes(n) + sampling_buffer_bytes(1ULL << n, sample_count, true);
    r->margin = qm_sampling_margin(gate_count);
    if (r->bytes > s->budget) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: a %u-qubit request (%llu bytes with its sampling buffers) exceeds the budget", what, n,
                  r->bytes);
        return -1;
    }
    r->gates.assign(gates, gates + gate_count);
    r->mats.resize(gate_count);
    for (UINT i = 0; i < gate_count; ++i) {
        const qm_batch_gate& g = gates[i];
        if (g.matrix && g.qubit_count >= 1 && g.qubit_count <= 4) {
            const size_t d = (size_t)1 << g.qubit_count;
            const size_t len = g.kind == QM_BATCH_DIAG ? d : d * d;
            r->mats[i].assign(g.matrix, g.matrix + len);
            r->gates[i].matrix = r->mats[i].data();
        }
    }
    r->bf.assign(bf, bf + term_count);
    r->pf.assign(pf, pf + term_count);
    r->rot.assign(rot, rot + term_count);
    r->coef.assign(coefs, coefs + term_count);
    r->uniforms.assign(uniforms, uniforms + sample_count);
    for (UINT t = 0; t < term_count; ++t)
        if (((bf[t] | pf[t]) >> n) != 0) {
            set_error(QM_ERR_INVALID_ARGUMENT, "%s: term %u acts beyond %u qubits", what, t, n);
            return -1;
        }
    if (r->small) {
        // validate on a host-only batch (no device work)
        qm_batch* probe = qm_batch_create();
        const int idx = qm_batch_add_circuit(probe, n, r->gates.data(), gate_count, bf, pf, rot, coefs, term_count);
        qm_batch_destroy(probe);
        if (idx < 0) return -1;
    } else {
        r->circuit = qm_circuit_create(n, r->gates.data(), gate_count, 0, 0);
        // the program goes to the device here, in the submitting thread, so the
        // scheduler thread never waits on an upload
        if (!r->circuit || !use_device(s->device) || !circuit_prepare(r->circuit, s->device)) return -1;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: server is shutting down", what);
        return -1;
    }
    r->id = s->next_id++;
    r->t_submit = clk::now();
    s->table[r->id] = r;
    s->inbox.push_back(r);
    s->cv_work.notify_one();
    return r->id;
}

int qm_server_wait(qm_server* s, long long id, double* energy, ITYPE* samples, double* latency_seconds) {
    return qm_server_wait_checked(s, id, energy, samples, nullptr, nullptr, latency_seconds);
}

int qm_server_wait_checked(qm_server* s, long long id, double* energy, ITYPE* samples, unsigned char* flags,
                           UINT* flagged, double* latency_seconds) {
    const char* what = "qm_server_wait";
    if (flagged) *flagged = 0;
    if (!s) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: null server", what);
        return QM_ERR_INVALID_ARGUMENT;
    }
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->table.find(id);
    if (it == s->table.end()) {
        set_error(QM_ERR_INVALID_ARGUMENT, "%s: unknown or already collected request %lld", what, id);
        return QM_ERR_INVALID_ARGUMENT;
    }
    ReqPtr r = it->second;
    s->cv_done.wait(lk, [&] { return r->status >= 2; });
    s->table.erase(id);
    lk.unlock();
    if (latency_seconds) *latency_seconds = seconds(r->t_submit, r->t_done);
    if (r->status == 3) {
        set_error(QM_ERR_HIP, "%s: request %lld failed: %s", what, id, r->error.c_str());
        return QM_ERR_HIP;
    }
    if (energy) *energy = r->energy;
    if (samples && !r->samples.empty()) std::memcpy(samples, r->samples.data(), r->samples.size() * 8);
    if (flags && !r->flags.empty()) std::memcpy(flags, r->flags.data(), r->flags.size());
    if (flagged) {
        UINT k = 0;
        for (unsigned char f : r->flags) k += f ? 1u : 0u;
        *flagged = k;
    }
    return QM_OK;
}

void qm_server_stats(qm_server* s, ITYPE* admitted, ITYPE* completed, ITYPE* failed, ITYPE* batches,
                     ITYPE* max_batch, ITYPE* peak_bytes) {
    if (!s) return;
    std::lock_guard<std::mutex> lk(s->mu);
    if (admitted) *admitted = s->n_admitted;
    if (completed) *completed = s->n_done;
    if (failed) *failed = s->n_failed;
    if (batches) *batches = s->n_batches;
    if (max_batch) *max_batch = s->max_batch;
    if (peak_bytes) *peak_bytes = s->peak_bytes;
}

void qm_server_destroy(qm_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice((int)s->device);
    if (s->bstream) (void)hipStreamSynchronize(s->bstream);
    for (Slot& sl : s->slots) {
        if (sl.sptr) release_cuda_stream_host(sl.sptr, 1, s->device);
        if (sl.ev) (void)hipEventDestroy(sl.ev);
        sl.energy.release();
        sl.samples.release();
        sl.uniforms.release();
        sl.flags.release();
    }
    if (s->batch) qm_batch_destroy(s->batch);
    if (s->bev) (void)hipEventDestroy(s->bev);
    if (s->bsptr) release_cuda_stream_host(s->bsptr, 1, s->device);
    s->benergy.release();
    s->bsamples.release();
    s->buniforms.release();
    s->bflags.release();
    delete s;
}

}  // extern "C"
