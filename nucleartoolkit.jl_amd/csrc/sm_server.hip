// Persistent serving loop for many main_sm eigenvalue solves on one device.
//
// Life of a request:
//   submit  -> prep queue (host threads: readsmsnt, basis, tables; the
//              interaction-independent index is built once per (orbits, Z_v,
//              N_v, 2M) and cached, later requests only build coefficients)
//           -> ready list, in arrival order
//           -> admitted to an idle slot when the HBM it needs fits the budget
//              (index tables if not resident + coefficient and workspace growth
//              of that slot); a later request may pass a waiting head request a
//              bounded number of times, so a large request is never starved
//           -> the slot's thread uploads the index if needed, binds the
//              coefficients into its own buffers and runs the TRL on its own
//              stream; solves of different slots overlap on the device
//           -> done / failed / canceled, latency stamped on the server clock.
// Slots keep their buffers between requests (grow-only), so a steady stream
// of similar requests allocates nothing; memory is reclaimed from idle slots
// and unused index tables only when an admission needs it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/smgpu.h"
#include "sm_device.h"
#include "sm_host.h"
#include "sm_status.h"
#include "sm_trl.h"

namespace smgpu {
namespace {

using Clock = std::chrono::steady_clock;

// Interaction-independent part of a request, shared by every request with the
// same orbits, valence numbers and 2M.
struct IndexEntry {
    std::string key;
    // host part (filled once by the first preparer; `ready` under the server mutex)
    bool ready = false;
    int32_t code = SMGPU_OK;
    std::string error;
    Basis basis;
    OneBodyOps ob[2];
    std::shared_ptr<const HostIndex> hidx;
    int64_t index_bytes = 0;
    // device part (server mutex)
    std::shared_ptr<DeviceIndex> didx;
    bool uploading = false;   // an admitted request is uploading it
    bool upload_failed = false;
    int users = 0;            // admitted requests holding it
    uint64_t last_use = 0;
};

struct Request {
    int64_t id = 0;
    std::string snt_path, nucleus;
    int twoM = -1;
    SolveParams params{};
    std::atomic<bool> cancel{false};
    int state = SMGPU_REQ_QUEUED;
    int32_t status = SMGPU_OK;
    std::string error;
    std::shared_ptr<IndexEntry> entry;
    std::shared_ptr<const HostCoefficients> coef;
    uint64_t coef_hash = 0;   // content_hash(*coef): candidates for a shared batch
    int64_t mdim = 0, ws_need = 0, coef_need = 0;
    bool uploader = false, index_shared = false;
    bool own_coef = false;    // a batch member whose coefficients differ from the run's
    int bypassed = 0;
    int slot = -1;
    int64_t reserved = 0;     // HBM charged at admission
    double t_submit = 0, t_prep0 = 0, t_prep1 = 0, t_admit = 0, t_done = 0;
    std::vector<double> evals;
    SolveStats st{};
};

// a slot's streams: its run stream, the batch H·v stream and those of members 1..
struct StreamSet {
    hipStream_t run = nullptr, hv = nullptr;
    std::vector<hipStream_t> members;
    void synchronize() const {
        for (hipStream_t x : all()) SM_HIP(hipStreamSynchronize(x));
    }
    void destroy() {
        for (hipStream_t x : all()) (void)hipStreamDestroy(x);
        run = hv = nullptr;
        members.clear();
    }
    std::vector<hipStream_t> all() const {
        std::vector<hipStream_t> v;
        if (run) v.push_back(run);
        if (hv) v.push_back(hv);
        for (hipStream_t x : members) if (x) v.push_back(x);
        return v;
    }
};

struct Slot {
    int id = 0;
    // sets[0]: the slot's streams (slot 0 of a CU partition: on the lane's CUs);
    // sets[1]: slot 0 of a CU partition only, the same on the whole device
    StreamSet sets[2];
    int bound_set = 0;
    hipStream_t stream = nullptr;
    SolveCtx ctx;             // one request, or member 0 of a batch
    // batches: the H·v stream and members 1.. (each its own stream and workspace)
    hipStream_t hv_stream = nullptr;
    std::vector<std::unique_ptr<SolveCtx>> more;
    DeviceHam ham;
    // members 1.. whose interaction differs from the run's: their own coefficients on
    // the run's index (grow-only buffers like ham's; member b uses mham[b - 1])
    std::vector<std::unique_ptr<DeviceHam>> mham;
    std::shared_ptr<Request> job;                  // first (or only) request
    // request per member slot of the run (null: free), sized at admission; changed
    // only under the server mutex (other threads count a slot's requests)
    std::vector<std::shared_ptr<Request>> batch;
    int admitted = 0;                              // requests placed at admission
    std::shared_ptr<const HostCoefficients> bound; // the run's Hamiltonian (for requests joining it)
    int64_t cap = 0;          // HBM held by this slot's buffers (server mutex)
    SolveCtx& member(int b) { return b == 0 ? ctx : *more[b - 1]; }
    int64_t workspace_bytes() const {
        int64_t c = workspace_capacity(ctx);
        for (auto& x : more) c += workspace_capacity(*x);
        return c;
    }
    void release_workspaces() {
        release_workspace(ctx);
        for (auto& x : more) release_workspace(*x);
    }
    int64_t coefficient_bytes() const {
        int64_t c = ham.coefficient_capacity();
        for (auto& h : mham) c += h->coefficient_capacity();
        return c;
    }
    void release_coefficients() {
        ham.release_coefficients();
        for (auto& h : mham) h->release_coefficients();
    }
    void drop_index() {
        ham.idx.reset();
        for (auto& h : mham) h->idx.reset();
    }
    // runs the next job on stream set k, once the work queued on the current set has drained
    void use_streams(int k) {
        if (k == bound_set || !sets[k].run) return;
        sets[bound_set].synchronize();
        const StreamSet& n = sets[k];
        stream = n.run;
        ctx.stream = n.run;
        hv_stream = n.hv;
        for (size_t b = 0; b < more.size(); b++) more[b]->stream = n.members[b];
        bound_set = k;
    }
    bool release = false;     // free the buffers before waiting again
    std::condition_variable cv;
    std::thread th;
};

}  // namespace

struct Server {
    smgpu_server_config cfg{};
    int64_t large_mdim = 0;   // > 0: large-lane admission order (schedule_priority_locked)
    int reserve_cus = 0;      // > 0: slot 0 is the large lane, on CUs reserve_cus..
    bool lane_split = false;  // slot 0 takes the large requests only, the other slots the rest
    Clock::time_point t0 = Clock::now();
    std::mutex m;
    std::condition_variable done_cv, prep_cv;
    bool stopping = false;
    int64_t next_id = 1;
    std::unordered_map<int64_t, std::shared_ptr<Request>> reqs;
    std::deque<std::shared_ptr<Request>> prep_q;
    std::list<std::shared_ptr<Request>> ready;
    std::map<std::string, std::shared_ptr<IndexEntry>> cache;
    std::condition_variable entry_cv;
    std::vector<std::unique_ptr<Slot>> slots;
    std::vector<std::thread> preparers;
    uint64_t use_clock = 0;
    // accounting
    int64_t index_resident = 0;   // bytes of resident or uploading index tables
    int64_t hbm_peak = 0;
    smgpu_server_stats st{};

    double now_ms() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }

    int64_t in_use() const {
        int64_t u = index_resident;
        for (auto& s : slots) u += s->cap;
        return u;
    }

    // ------------------------------------------------------------ preparation
    static std::string key_of(const Snt& s, const NucleusSpec& ns) {
        std::ostringstream k;
        for (const auto& o : s.p_sps) k << o.n << ',' << o.l << ',' << o.j2 << ',' << o.tz << ';';
        k << '|';
        for (const auto& o : s.n_sps) k << o.n << ',' << o.l << ',' << o.j2 << ',' << o.tz << ';';
        k << '|' << ns.vp << ',' << ns.vn << ',' << ns.twoM;
        return k.str();
    }

    // The index tables are cached per (orbits, Z_v, N_v, 2M).  The pp / nn tables
    // keep only operators with a nonzero coefficient, so their sparsity depends on
    // the interaction: a request whose pattern differs from the cached entry's gets
    // an entry of its own, keyed by the basis and its pattern.
    void prepare(Request& r) {
        const int nthr = std::max(1, cfg.table_threads);
        Snt snt = read_snt(r.snt_path, r.nucleus, false);
        NucleusSpec ns = resolve_nucleus(snt, r.nucleus, r.twoM);
        const std::string base_key = key_of(snt, ns);
        std::string key = base_key;
        std::shared_ptr<IndexEntry> e;
        for (int pass = 0;; pass++) {
            bool builder = false;
            {
                std::unique_lock<std::mutex> lk(m);
                auto it = cache.find(key);
                if (it == cache.end()) {
                    e = std::make_shared<IndexEntry>();
                    e->key = key;
                    cache[key] = e;
                    builder = true;
                    st.index_misses++;
                } else {
                    e = it->second;
                    st.index_hits++;
                    entry_cv.wait(lk, [&] { return e->ready; });
                    if (e->code != SMGPU_OK) throw std::invalid_argument("index tables failed: " + e->error);
                }
            }
            if (builder) {
                try {
                    Basis b = build_basis(snt, ns.vp, ns.vn, ns.twoM);
                    if (b.mdim == 0) throw std::invalid_argument("empty M-scheme space (mdim = 0)");
                    Tables t = build_tables(snt, b, nthr);
                    auto hidx = pack_index(b, t);
                    auto coef = pack_coefficients(b, t);
                    std::lock_guard<std::mutex> lk(m);
                    e->ob[0] = t.ob[0];
                    e->ob[1] = t.ob[1];
                    e->basis = std::move(b);
                    e->hidx = hidx;
                    e->index_bytes = device_bytes(*hidx);
                    e->ready = true;
                    r.coef = coef;
                } catch (...) {
                    std::string msg;
                    int32_t c = status_of_current_exception(msg);
                    std::lock_guard<std::mutex> lk(m);
                    e->code = c;
                    e->error = msg;
                    e->ready = true;
                    cache.erase(key);   // a later request retries
                    entry_cv.notify_all();
                    throw;
                }
                entry_cv.notify_all();
                break;
            }
            Tables t;
            t.ob[0] = e->ob[0];
            t.ob[1] = e->ob[1];
            build_coefficients(snt, e->basis, t, nthr);
            r.coef = pack_coefficients(e->basis, t);
            if (pattern_of(*r.coef) == pattern_of(*e->hidx)) break;
            if (pass > 0) throw std::logic_error("index entry keyed by a sparsity pattern holds another pattern");
            std::ostringstream k2;
            k2 << base_key << "#pattern=" << std::hex << pattern_of(*r.coef);
            key = k2.str();
            r.coef.reset();
        }
        r.entry = e;
        r.coef_hash = content_hash(*r.coef);
        r.mdim = e->basis.mdim;
        r.ws_need = solve_workspace_bytes(r.mdim, r.params);
        r.coef_need = device_bytes(*r.coef);
        bound_basis(r, e->index_bytes);
    }

    // A request whose Krylov basis at its lm does not fit the whole HBM budget (with
    // its index tables and coefficients) runs with the largest lm that does, down to
    // max(ls + 20, 2 ls); the thick restart then converges to the same eigenvalues
    // in more sweeps.  The bound depends on the request and the budget only, never on
    // what else is running, so a request gives the same result alone and in a mix.
    // Below that floor the request keeps its lm and admission fails it with ENOMEM.
    void bound_basis(Request& r, int64_t index_bytes) {
        const int64_t fixed = index_bytes + r.coef_need;
        if (fixed + r.ws_need <= cfg.hbm_budget) return;
        const int lm0 = r.params.lm;
        const int lm_min = std::min(lm0, std::max(r.params.ls + 20, 2 * r.params.ls));
        int lo = lm_min, hi = lm0 - 1, best = 0;
        while (lo <= hi) {
            const int mid = lo + (hi - lo) / 2;
            SolveParams q = r.params;
            q.lm = mid;
            if (fixed + solve_workspace_bytes(r.mdim, q) <= cfg.hbm_budget) { best = mid; lo = mid + 1; }
            else hi = mid - 1;
        }
        if (best == 0) return;
        r.params.lm = best;
        r.ws_need = solve_workspace_bytes(r.mdim, r.params);
    }

    void preparer_loop() {
        for (;;) {
            std::shared_ptr<Request> r;
            {
                std::unique_lock<std::mutex> lk(m);
                prep_cv.wait(lk, [&] { return stopping || !prep_q.empty(); });
                if (stopping) return;
                r = prep_q.front();
                prep_q.pop_front();
                if (r->cancel) { finish_locked(*r, SMGPU_REQ_CANCELED, SMGPU_ECANCELED, "canceled"); continue; }
                r->state = SMGPU_REQ_PREPARING;
                r->t_prep0 = now_ms();
            }
            int32_t code = SMGPU_OK;
            std::string msg;
            try {
                prepare(*r);
            } catch (...) {
                code = status_of_current_exception(msg);
            }
            std::lock_guard<std::mutex> lk(m);
            r->t_prep1 = now_ms();
            if (code != SMGPU_OK) { finish_locked(*r, SMGPU_REQ_FAILED, code, msg); continue; }
            if (r->cancel) { finish_locked(*r, SMGPU_REQ_CANCELED, SMGPU_ECANCELED, "canceled"); continue; }
            r->state = SMGPU_REQ_READY;
            ready.push_back(r);
            schedule_locked();
        }
    }

    // ------------------------------------------------------------- admission
    void finish_locked(Request& r, int state, int32_t code, const std::string& msg) {
        r.state = state;
        r.status = code;
        r.error = msg;
        r.t_done = now_ms();
        r.coef.reset();
        if (state == SMGPU_REQ_DONE) st.completed++;
        else if (state == SMGPU_REQ_CANCELED) st.canceled++;
        else st.failed++;
        done_cv.notify_all();
    }

    // free HBM not held by running work: unused index tables (least recently
    // used first) and the buffers of idle slots; true if anything was freed
    bool reclaim_locked(const IndexEntry* keep) {
        std::vector<IndexEntry*> cand;
        for (auto& kv : cache) {
            IndexEntry* e = kv.second.get();
            if (e != keep && e->didx && e->users == 0 && !e->uploading) cand.push_back(e);
        }
        std::sort(cand.begin(), cand.end(), [](IndexEntry* a, IndexEntry* b) { return a->last_use < b->last_use; });
        if (!cand.empty()) {
            IndexEntry* e = cand.front();
            index_resident -= e->index_bytes;
            e->didx.reset();
            st.index_evictions++;
            return true;
        }
        for (auto& s : slots)
            if (!s->job && s->cap > 0) {
                s->drop_index();
                s->release_coefficients();
                s->release_workspaces();
                s->cap = 0;
                return true;
            }
        return false;
    }

    bool try_admit_locked(const std::shared_ptr<Request>& r) {
        IndexEntry& e = *r->entry;
        const int64_t need_index = (e.didx || e.uploading) ? 0 : e.index_bytes;
        if (need_index + r->ws_need + r->coef_need > cfg.hbm_budget) {
            ready.remove(r);
            std::ostringstream os;
            os << "does not fit the HBM budget: needs " << (need_index + r->ws_need + r->coef_need) << " bytes, budget "
               << cfg.hbm_budget;
            finish_locked(*r, SMGPU_REQ_FAILED, SMGPU_ENOMEM, os.str());
            return true;
        }
        // requests that can share this one's H·v launches: same index tables and
        // precision, identical coefficients (or any interaction on the basis with
        // batch_variants; those bring their own coefficients), in ready-list order
        std::vector<std::shared_ptr<Request>> mates;
        std::vector<char> own;    // mate i needs its own coefficient buffers
        if (cfg.max_batch > 1)
            for (auto& q : ready) {
                if ((int)mates.size() + 1 >= cfg.max_batch) break;
                if (q == r || q->cancel || q->entry != r->entry) continue;
                if ((q->params.precision == 0) != (r->params.precision == 0)) continue;
                const bool same = q->coef_hash == r->coef_hash && same_coefficients(*q->coef, *r->coef);
                if (!same && !cfg.batch_variants) continue;
                mates.push_back(q);
                own.push_back(same ? 0 : 1);
            }
        for (;;) {
            Slot* best = nullptr;
            int64_t best_g = 0;
            int best_n = 1;
            for (auto& s : slots) {
                if (s->job) continue;
                if (lane_split && (r->mdim >= large_mdim) != (s->id == 0)) continue;
                // the largest batch (up to the mates found) whose growth fits the budget
                int64_t g = coefficient_growth(s->ham, *r->coef) + workspace_growth(s->ctx, r->mdim, r->params);
                int n = 1;
                for (size_t i = 0; i < mates.size(); i++) {
                    int64_t gi = workspace_growth(s->member((int)i + 1), mates[i]->mdim, mates[i]->params);
                    if (own[i]) gi += coefficient_growth(*s->mham[i + 1], *mates[i]->coef);
                    if (in_use() + need_index + g + gi > cfg.hbm_budget) break;
                    g += gi;
                    n++;
                }
                if (!best || n > best_n || (n == best_n && g < best_g)) { best = s.get(); best_g = g; best_n = n; }
            }
            if (!best) return false;
            if (in_use() + need_index + best_g <= cfg.hbm_budget) {
                std::vector<std::shared_ptr<Request>> run{r};
                r->own_coef = false;
                for (int i = 0; i + 1 < best_n; i++) {
                    run.push_back(mates[i]);
                    mates[i]->own_coef = own[i] != 0;
                }
                for (auto& q : run) {
                    ready.remove(q);
                    q->state = SMGPU_REQ_RUNNING;
                    q->slot = best->id;
                    q->t_admit = now_ms();
                    q->index_shared = e.didx != nullptr || e.uploading;
                    e.users++;
                }
                if (need_index) { e.uploading = true; r->uploader = true; index_resident += e.index_bytes; }
                e.last_use = ++use_clock;
                r->reserved = need_index + best_g;
                best->cap += best_g;
                best->job = r;
                best->batch = run;
                best->admitted = (int)run.size();
                best->batch.resize(std::max<size_t>(run.size(), (size_t)cfg.max_batch));
                if (run.size() > 1) {
                    st.batched += (int64_t)run.size();
                    st.peak_batch = std::max(st.peak_batch, (int32_t)run.size());
                }
                hbm_peak = std::max(hbm_peak, in_use());
                int running = 0;
                for (auto& s : slots) running += running_in(*s);
                st.peak_running = std::max(st.peak_running, running);
                best->cv.notify_one();
                return true;
            }
            if (!reclaim_locked(&e)) return false;
        }
    }

    bool large_running() const {
        for (auto& s : slots)
            if (s->job && s->job->mdim >= large_mdim) return true;
        return false;
    }

    // Large-lane order: requests of at least large_mdim run one solve (batch) at a
    // time, oldest first, as soon as none holds a slot; the others, oldest first once
    // they have waited 60 s, else smallest basis first.  No request holds up another.
    void schedule_priority_locked() {
        constexpr double AGE_MS = 60000.0;
        for (bool admitted = true; admitted;) {
            admitted = false;
            const double now = now_ms();
            const bool lane_busy = large_running();
            std::vector<std::shared_ptr<Request>> order(ready.begin(), ready.end());
            auto rank = [&](const std::shared_ptr<Request>& r) {
                if (r->mdim >= large_mdim) return 0;
                return now - r->t_submit > AGE_MS ? 1 : 2;
            };
            std::stable_sort(order.begin(), order.end(), [&](const auto& a, const auto& b) {
                const int ra = rank(a), rb = rank(b);
                if (ra != rb) return ra < rb;
                if (ra == 2 && a->mdim != b->mdim) return a->mdim < b->mdim;
                return a->t_submit < b->t_submit;
            });
            for (auto& r : order) {
                if (r->mdim >= large_mdim && lane_busy) continue;
                if (try_admit_locked(r)) { admitted = true; break; }
            }
        }
    }

    void schedule_locked() {
        if (stopping) return;
        if (large_mdim > 0) { schedule_priority_locked(); return; }
        const int max_bypass = 4 * (int)slots.size();
        // An admission removes the request and any batch mates from the ready list,
        // so every admission ends the pass and the scan starts again from the front.
        for (bool admitted = true; admitted;) {
            admitted = false;
            bool head_blocked = false;
            for (auto it = ready.begin(); it != ready.end(); ++it) {
                if (head_blocked && ready.front()->bypassed >= max_bypass) break;
                auto r = *it;
                if (try_admit_locked(r)) {
                    if (head_blocked && !ready.empty() && r != ready.front()) ready.front()->bypassed++;
                    admitted = true;
                    break;
                }
                head_blocked = true;
            }
        }
    }

    // ---------------------------------------------------------------- slots
    void run_job(Slot& s, Request& r) {
        IndexEntry& e = *r.entry;
        std::shared_ptr<DeviceIndex> idx;
        if (r.uploader) {
            std::shared_ptr<DeviceIndex> up;
            try {
                up = upload_index(*e.hidx, cfg.device);
            } catch (...) {
                std::lock_guard<std::mutex> lk(m);
                e.uploading = false;
                e.upload_failed = true;
                index_resident -= e.index_bytes;
                entry_cv.notify_all();
                throw;
            }
            std::lock_guard<std::mutex> lk(m);
            e.didx = up;
            e.uploading = false;
            e.upload_failed = false;
            st.index_uploads++;
            entry_cv.notify_all();
            idx = up;
        } else {
            std::unique_lock<std::mutex> lk(m);
            entry_cv.wait(lk, [&] { return !e.uploading; });
            if (!e.didx) throw std::runtime_error("index upload failed in another request");
            idx = e.didx;
        }
        bind_ham(s.ham, idx, *r.coef);
        const int nb = s.admitted;
        for (int b = 0; b < nb; b++) s.batch[b]->evals.assign(std::max(1, s.batch[b]->params.num_ev), 0.0);
        const int cap = (int)s.batch.size();
        struct Unhook {
            Slot& s;
            int cap;
            ~Unhook() {
                for (int b = 0; b < cap; b++) s.member(b).cancel = nullptr;
                s.bound.reset();
            }
        } unhook{s, cap};
        for (int b = 0; b < nb; b++) s.member(b).cancel = &s.batch[b]->cancel;
        if (cap == 1) {
            trl_solve(s.ham, r.params, r.evals.data(), nullptr, r.st, nullptr, &s.ctx);
            r.evals.resize((size_t)std::min<int64_t>(r.params.num_ev, r.mdim));
            return;
        }
        // lockstep batch with member slots; later requests on the same basis (the same
        // Hamiltonian, or any interaction with batch_variants) join free member slots
        // between rounds; each result is published as soon as its solve finishes.
        // Members with the run's coefficients use s.ham, the others their slot's mham.
        s.bound = r.coef;
        const uint64_t hash = r.coef_hash;
        const bool fp64 = r.params.precision == 0;
        std::vector<DeviceHam*> mh(cap, &s.ham);
        for (int b = 1; b < nb; b++)
            if (s.batch[b]->own_coef) {
                bind_ham(*s.mham[b], idx, *s.batch[b]->coef);
                mh[b] = s.mham[b].get();
            }
        std::vector<SolveParams> p(cap);
        std::vector<double*> ev(cap, nullptr);
        std::vector<SolveStats> sst(cap);
        std::vector<const SolveCtx*> cx(cap);
        std::vector<int32_t> status(cap, SMGPU_OK);
        std::vector<std::string> msgs(cap);
        auto fill = [&](int b) {
            Request& q = *s.batch[b];
            p[b] = q.params;
            p[b].num_ev = (int32_t)std::min<int64_t>(p[b].num_ev, q.mdim);
            ev[b] = q.evals.data();
            sst[b] = SolveStats{};
            status[b] = SMGPU_OK;
            msgs[b].clear();
        };
        for (int b = 0; b < cap; b++) cx[b] = &s.member(b);
        for (int b = 0; b < nb; b++) fill(b);
        auto join = [&](int b) -> bool {
            std::shared_ptr<Request> keep;
            {
            std::lock_guard<std::mutex> lk(m);
            if (stopping) return false;
            IndexEntry& e = *r.entry;
            for (auto& q : ready) {
                if (q->cancel || q->entry != r.entry || (q->params.precision == 0) != fp64) continue;
                const bool same = q->coef_hash == hash && same_coefficients(*q->coef, *s.bound);
                if (!same && !cfg.batch_variants) continue;
                int64_t g = workspace_growth(s.member(b), q->mdim, q->params);
                if (!same) g += coefficient_growth(*s.mham[b], *q->coef);
                if (in_use() + g > cfg.hbm_budget && !(reclaim_locked(&e) && in_use() + g <= cfg.hbm_budget))
                    return false;
                keep = q;               // ready.remove drops the list's reference
                keep->own_coef = !same;
                ready.remove(keep);
                keep->state = SMGPU_REQ_RUNNING;
                keep->slot = s.id;
                keep->t_admit = now_ms();
                keep->index_shared = true;
                keep->reserved = g;
                e.users++;
                e.last_use = ++use_clock;
                s.cap += g;
                hbm_peak = std::max(hbm_peak, in_use());
                s.batch[b] = keep;
                s.member(b).cancel = &keep->cancel;
                st.batched++;
                int live = 0;
                for (auto& x : s.batch) live += (x && !finished(*x)) ? 1 : 0;
                st.peak_batch = std::max(st.peak_batch, (int32_t)live);
                int running = 0;
                for (auto& sl : slots) running += running_in(*sl);
                st.peak_running = std::max(st.peak_running, running);
                break;
            }
            if (!keep) return false;
            }
            // the upload of a variant's coefficients runs outside the server mutex
            mh[b] = &s.ham;
            if (keep->own_coef) {
                try {
                    bind_ham(*s.mham[b], idx, *keep->coef);
                } catch (...) {
                    std::string msg;
                    const int32_t code = status_of_current_exception(msg);
                    std::lock_guard<std::mutex> lk(m);
                    keep->entry->users--;
                    s.member(b).cancel = nullptr;
                    s.batch[b].reset();
                    finish_locked(*keep, SMGPU_REQ_FAILED, code, msg);
                    return false;
                }
                mh[b] = s.mham[b].get();
            }
            keep->evals.assign(std::max(1, keep->params.num_ev), 0.0);
            fill(b);
            return true;
        };
        BatchIO io;
        io.hams = mh.data();
        io.p = p.data();
        io.evals = ev.data();
        io.st = sst.data();
        io.status = status.data();
        io.msgs = msgs.data();
        trl_solve_batch(s.ham, cap, nb, io, cx.data(), s.hv_stream, join, [&](int b) {
            Request& q = *s.batch[b];
            q.st = sst[b];
            q.evals.resize((size_t)p[b].num_ev);
            std::lock_guard<std::mutex> lk(m);
            q.entry->users--;
            q.entry->last_use = ++use_clock;
            s.member(b).cancel = nullptr;
            if (status[b] == SMGPU_OK) finish_locked(q, SMGPU_REQ_DONE, SMGPU_OK, "");
            else if (status[b] == SMGPU_ECANCELED) finish_locked(q, SMGPU_REQ_CANCELED, status[b], msgs[b]);
            else finish_locked(q, SMGPU_REQ_FAILED, status[b], msgs[b]);
        });
    }

    // CU partition: whether a request below large_mdim is queued, being prepared (its size
    // still unknown), ready or running, so that the large lane keeps to its CUs (server mutex)
    bool smaller_work_locked() const {
        for (const auto& kv : reqs) {
            const Request& q = *kv.second;
            if (q.state == SMGPU_REQ_QUEUED || q.state == SMGPU_REQ_PREPARING) return true;
            if (q.state == SMGPU_REQ_READY && q.mdim < large_mdim) return true;
            if (q.state == SMGPU_REQ_RUNNING && q.slot != 0) return true;
        }
        return false;
    }

    // requests running in a slot (server mutex)
    static int running_in(const Slot& s) {
        if (!s.job) return 0;
        int n = 0;
        for (auto& x : s.batch) n += (x && !finished(*x)) ? 1 : 0;
        return std::max(n, 1);
    }

    void slot_loop(Slot& s) {
        (void)hipSetDevice(cfg.device);
        for (;;) {
            std::shared_ptr<Request> r;
            int set = 0;
            {
                std::unique_lock<std::mutex> lk(m);
                s.cv.wait(lk, [&] { return stopping || s.job; });
                if (!s.job) return;   // stopping
                r = s.job;
                if (s.sets[1].run && !smaller_work_locked()) set = 1;
            }
            int32_t code = SMGPU_OK;
            std::string msg;
            try {
                s.use_streams(set);
                // (a batch member's own cancel flag is checked at its first step)
                if (r->cancel && cfg.max_batch == 1) throw Canceled();
                run_job(s, *r);
            } catch (...) {
                code = status_of_current_exception(msg);
            }
            std::lock_guard<std::mutex> lk(m);
            s.drop_index();
            // requests the batch driver has not finished (one request alone, or a run
            // that failed before the driver started)
            for (size_t b = 1; b < s.batch.size(); b++) {
                if (!s.batch[b]) continue;
                Request& q = *s.batch[b];
                if (finished(q)) continue;
                q.entry->users--;
                if (code == SMGPU_ECANCELED || q.cancel) finish_locked(q, SMGPU_REQ_CANCELED, SMGPU_ECANCELED, "canceled");
                else finish_locked(q, SMGPU_REQ_FAILED, code == SMGPU_OK ? SMGPU_EINVAL : code, msg);
            }
            const bool open = !finished(*r);
            if (open) {
                r->entry->users--;
                r->entry->last_use = ++use_clock;
            }
            // charge what the slot really holds now
            s.cap = s.workspace_bytes() + s.coefficient_bytes();
            if (s.cap > cfg.keep_bytes) {
                s.release_coefficients();
                s.release_workspaces();
                s.cap = 0;
            }
            s.job.reset();
            s.batch.clear();
            if (!open) {
            } else if (code == SMGPU_OK) finish_locked(*r, SMGPU_REQ_DONE, SMGPU_OK, "");
            else if (code == SMGPU_ECANCELED) finish_locked(*r, SMGPU_REQ_CANCELED, code, msg);
            else finish_locked(*r, SMGPU_REQ_FAILED, code, msg);
            schedule_locked();
        }
    }

    // ------------------------------------------------------------------ API
    explicit Server(const smgpu_server_config& c) : cfg(c) {
        SM_HIP(hipSetDevice(cfg.device));
        if (cfg.max_inflight <= 0) cfg.max_inflight = 4;
        if (cfg.max_inflight > 64) throw std::invalid_argument("max_inflight above 64");
        if (cfg.host_threads <= 0) cfg.host_threads = 2;
        if (cfg.table_threads <= 0)
            cfg.table_threads = std::max(1, (int)std::thread::hardware_concurrency() / cfg.host_threads);
        if (cfg.keep_bytes <= 0) cfg.keep_bytes = (int64_t)16 << 30;
        if (cfg.max_batch <= 0) cfg.max_batch = 1;
        if (cfg.max_batch > 16) throw std::invalid_argument("max_batch above 16");
        if (cfg.large_mdim < 0 || cfg.reserve_cus < 0) throw std::invalid_argument("negative large_mdim or reserve_cus");
        if (cfg.reserve_cus > 0 && (cfg.large_mdim == 0 || cfg.max_inflight < 2))
            throw std::invalid_argument("reserve_cus needs large_mdim > 0 and max_inflight >= 2");
        large_mdim = cfg.large_mdim;
        reserve_cus = cfg.reserve_cus;
        if (cfg.hbm_budget <= 0) {
            size_t fr = 0, tot = 0;
            SM_HIP(hipMemGetInfo(&fr, &tot));
            cfg.hbm_budget = (int64_t)(0.9 * (double)fr);
        }
        // large lane on a CU partition: slot 0's streams leave the first reserve_cus CUs
        // to the other slots, whose short kernels then never queue behind a large H·v
        std::vector<uint32_t> lane_mask;
        if (reserve_cus > 0) {
            hipDeviceProp_t prop{};
            SM_HIP(hipGetDeviceProperties(&prop, cfg.device));
            const int ncu = prop.multiProcessorCount;
            if (reserve_cus > ncu / 2) throw std::invalid_argument("reserved CUs above half the device");
            lane_mask.assign((size_t)(ncu + 31) / 32, 0u);
            for (int c = reserve_cus; c < ncu; c++) lane_mask[(size_t)c / 32] |= 1u << (c % 32);
        }
        auto new_stream = [&](hipStream_t* st, bool lane) {
            if (lane && !lane_mask.empty())
                SM_HIP(hipExtStreamCreateWithCUMask(st, (uint32_t)lane_mask.size(), lane_mask.data()));
            else
                SM_HIP(hipStreamCreateWithFlags(st, hipStreamNonBlocking));
        };
        auto new_set = [&](StreamSet& set, bool lane) {
            new_stream(&set.run, lane);
            if (cfg.max_batch > 1) {
                new_stream(&set.hv, lane);
                set.members.assign((size_t)cfg.max_batch - 1, nullptr);
                for (auto& x : set.members) new_stream(&x, lane);
            }
        };
        for (int i = 0; i < cfg.max_inflight; i++) {
            auto s = std::make_unique<Slot>();
            s->id = i;
            slots.push_back(std::move(s));
            Slot& sl = *slots.back();
            new_set(sl.sets[0], i == 0);
            // slot 0 of a CU partition also gets whole-device streams: a CU-masked queue
            // slows a large solve by about 14% even with 4 CUs reserved, so the lane runs
            // on the partition only while smaller requests are in the server
            if (i == 0 && !lane_mask.empty()) new_set(sl.sets[1], false);
            sl.stream = sl.sets[0].run;
            sl.ctx.stream = sl.stream;
            if (cfg.max_batch > 1) {
                for (int b = 0; b < cfg.max_batch; b++) sl.mham.push_back(std::make_unique<DeviceHam>());
                sl.hv_stream = sl.sets[0].hv;
                for (int b = 1; b < cfg.max_batch; b++) {
                    auto c = std::make_unique<SolveCtx>();
                    c->stream = sl.sets[0].members[(size_t)b - 1];
                    sl.more.push_back(std::move(c));
                }
            }
        }
        lane_split = !lane_mask.empty();
        for (auto& s : slots) {
            Slot* sp = s.get();
            sp->th = std::thread([this, sp] { slot_loop(*sp); });
        }
        for (int i = 0; i < cfg.host_threads; i++) preparers.emplace_back([this] { preparer_loop(); });
    }

    ~Server() {
        (void)hipSetDevice(cfg.device);
        {
            std::lock_guard<std::mutex> lk(m);
            stopping = true;
            for (auto& kv : reqs) kv.second->cancel = true;
            prep_cv.notify_all();
            for (auto& s : slots) s->cv.notify_all();
        }
        for (auto& t : preparers) t.join();
        for (auto& s : slots) s->th.join();
        std::lock_guard<std::mutex> lk(m);
        for (auto& r : prep_q) finish_locked(*r, SMGPU_REQ_CANCELED, SMGPU_ECANCELED, "server destroyed");
        for (auto& r : ready) finish_locked(*r, SMGPU_REQ_CANCELED, SMGPU_ECANCELED, "server destroyed");
        prep_q.clear();
        ready.clear();
        for (auto& s : slots) {
            s->drop_index();
            s->release_workspaces();
            s->sets[0].destroy();
            s->sets[1].destroy();
        }
        cache.clear();
    }

    int64_t submit(const char* path, const char* nuc, int twoM, const SolveParams& p) {
        if (p.num_ev < 1 || p.num_ev > 64) throw std::invalid_argument("num_ev must be in 1..64");
        auto r = std::make_shared<Request>();
        r->snt_path = path;
        r->nucleus = nuc;
        r->twoM = twoM;
        r->params = p;
        if (r->params.max_iter <= 0) r->params.max_iter = 100000;
        std::lock_guard<std::mutex> lk(m);
        if (stopping) throw std::invalid_argument("server is shutting down");
        r->id = next_id++;
        r->t_submit = now_ms();
        reqs[r->id] = r;
        prep_q.push_back(r);
        st.submitted++;
        prep_cv.notify_one();
        return r->id;
    }

    std::shared_ptr<Request> find_locked(int64_t id) {
        auto it = reqs.find(id);
        if (it == reqs.end()) throw std::invalid_argument("unknown request id");
        return it->second;
    }

    static bool finished(const Request& r) { return r.state >= SMGPU_REQ_DONE; }

    void fill_info(const Request& r, smgpu_request_info* info) {
        std::memset(info, 0, sizeof(*info));
        info->state = r.state;
        info->status = r.status;
        info->num_ev = (int32_t)r.evals.size();
        info->index_shared = r.index_shared ? 1 : 0;
        info->slot = r.slot;
    info->lm = r.params.lm;
        info->mdim = r.mdim;
        info->index_bytes = r.entry ? r.entry->index_bytes : 0;
        info->footprint = info->index_bytes + r.ws_need + r.coef_need;
        info->submit_ms = r.t_submit;
        info->prep_start_ms = r.t_prep0;
        info->prep_end_ms = r.t_prep1;
        info->admit_ms = r.t_admit;
        info->done_ms = r.t_done;
        info->latency_ms = finished(r) ? r.t_done - r.t_submit : 0.0;
        info->solve = r.st;
    }

    void cancel(int64_t id) {
        std::lock_guard<std::mutex> lk(m);
        auto r = find_locked(id);
        if (finished(*r)) return;
        r->cancel = true;
        if (r->state == SMGPU_REQ_QUEUED) {
            prep_q.erase(std::remove(prep_q.begin(), prep_q.end(), r), prep_q.end());
            finish_locked(*r, SMGPU_REQ_CANCELED, SMGPU_ECANCELED, "canceled");
        } else if (r->state == SMGPU_REQ_READY) {
            ready.remove(r);
            finish_locked(*r, SMGPU_REQ_CANCELED, SMGPU_ECANCELED, "canceled");
            schedule_locked();
        }
        // preparing: dropped when the preparation ends; running: the solve
        // polls the flag once per Lanczos iteration
    }

    void stats(smgpu_server_stats* out) {
        std::lock_guard<std::mutex> lk(m);
        smgpu_server_stats s = st;
        s.queued = (int32_t)prep_q.size();
        s.preparing = 0;
        s.ready = (int32_t)ready.size();
        s.running = 0;
        for (auto& kv : reqs) if (kv.second->state == SMGPU_REQ_PREPARING) s.preparing++;
        for (auto& sl : slots) s.running += running_in(*sl);
        s.slots = (int32_t)slots.size();
        s.hbm_budget = cfg.hbm_budget;
        s.hbm_in_use = in_use();
        s.hbm_peak = hbm_peak;
        s.clock_ms = now_ms();
        *out = s;
    }
};

}  // namespace smgpu

using namespace smgpu;

struct smgpu_server_s {
    std::unique_ptr<Server> impl;
};

extern "C" {

void smgpu_default_server_config(smgpu_server_config* cfg) {
    if (!cfg) return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->device = 0;
    cfg->max_inflight = 4;
    cfg->hbm_budget = 0;
    cfg->host_threads = 2;
    cfg->table_threads = 0;
    cfg->keep_bytes = 0;
    cfg->max_batch = 1;
    cfg->batch_variants = 1;
    cfg->large_mdim = 0;
    cfg->reserve_cus = 0;
}

int32_t smgpu_server_create(const smgpu_server_config* cfg, smgpu_server* out) {
    return guard([&]() -> int32_t {
        if (!out) throw std::invalid_argument("null argument");
        *out = nullptr;
        smgpu_server_config c;
        smgpu_default_server_config(&c);
        if (cfg) c = *cfg;
        auto s = std::make_unique<smgpu_server_s>();
        s->impl = std::make_unique<Server>(c);
        *out = s.release();
        return SMGPU_OK;
    });
}

int32_t smgpu_server_destroy(smgpu_server s) {
    return guard([&]() -> int32_t {
        delete s;
        return SMGPU_OK;
    });
}

int32_t smgpu_server_submit(smgpu_server s, const char* snt_path, const char* nucleus, int32_t twoM,
                            const smgpu_solve_params* params, int64_t* request_id) {
    return guard([&]() -> int32_t {
        if (!s || !snt_path || !nucleus || !params || !request_id) throw std::invalid_argument("null argument");
        *request_id = s->impl->submit(snt_path, nucleus, twoM, *params);
        return SMGPU_OK;
    });
}

int32_t smgpu_server_poll(smgpu_server s, int64_t request_id, int32_t* state) {
    return guard([&]() -> int32_t {
        if (!s || !state) throw std::invalid_argument("null argument");
        std::lock_guard<std::mutex> lk(s->impl->m);
        *state = s->impl->find_locked(request_id)->state;
        return SMGPU_OK;
    });
}

int32_t smgpu_server_wait(smgpu_server s, int64_t request_id, double timeout_ms) {
    return guard([&]() -> int32_t {
        if (!s) throw std::invalid_argument("null argument");
        Server& S = *s->impl;
        std::unique_lock<std::mutex> lk(S.m);
        auto r = S.find_locked(request_id);
        auto pred = [&] { return Server::finished(*r); };
        if (timeout_ms < 0) {
            S.done_cv.wait(lk, pred);
            return SMGPU_OK;
        }
        bool ok = S.done_cv.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), pred);
        if (!ok) set_last_error("timed out");
        return ok ? SMGPU_OK : SMGPU_ETIMEDOUT;
    });
}

int32_t smgpu_server_result(smgpu_server s, int64_t request_id, double* evals, smgpu_request_info* info) {
    return guard([&]() -> int32_t {
        if (!s) throw std::invalid_argument("null argument");
        Server& S = *s->impl;
        std::lock_guard<std::mutex> lk(S.m);
        auto r = S.find_locked(request_id);
        if (!Server::finished(*r)) throw std::invalid_argument("request has not finished");
        if (info) S.fill_info(*r, info);
        if (evals) std::copy(r->evals.begin(), r->evals.end(), evals);
        S.reqs.erase(request_id);
        if (r->status != SMGPU_OK) set_last_error(r->error);
        return r->status;
    });
}

int32_t smgpu_server_info(smgpu_server s, int64_t request_id, smgpu_request_info* info) {
    return guard([&]() -> int32_t {
        if (!s || !info) throw std::invalid_argument("null argument");
        std::lock_guard<std::mutex> lk(s->impl->m);
        s->impl->fill_info(*s->impl->find_locked(request_id), info);
        return SMGPU_OK;
    });
}

int32_t smgpu_server_cancel(smgpu_server s, int64_t request_id) {
    return guard([&]() -> int32_t {
        if (!s) throw std::invalid_argument("null argument");
        s->impl->cancel(request_id);
        return SMGPU_OK;
    });
}

int32_t smgpu_server_get_stats(smgpu_server s, smgpu_server_stats* st) {
    return guard([&]() -> int32_t {
        if (!s || !st) throw std::invalid_argument("null argument");
        s->impl->stats(st);
        return SMGPU_OK;
    });
}

}  // extern "C"
