// Continuous-batching request server (see serve.h).
#include "serve.h"

#include <algorithm>
#include <cstring>

namespace tdda {

Server::Server(const ServerOptions& o, const SiciDevice& tables) : opt_(o), tables_(tables) {
    TDDA_REQUIRE(o.max_slots >= 1 && o.max_slots <= 4096, ERR_INVALID, "max_slots must be in [1, 4096]");
    TDDA_REQUIRE(o.chunk >= 1, ERR_INVALID, "chunk must be positive");
    t0_ = std::chrono::steady_clock::now();
    TDDA_HIP(hipStreamCreate(&main_));
    TDDA_HIP(hipStreamCreate(&setup_));
    ks_.alloc(o.max_slots);
    ks_.zero(main_);
    for (int i = 0; i < 2; ++i) {   // rebuilt while the GPU runs the other batch
        batch_[i].set_pinned(true);
        init_batch_[i].set_pinned(true);
    }
    for (int i = 0; i < 2; ++i) {
        TDDA_HIP(hipHostMalloc((void**)&snap_[i], o.max_slots * sizeof(KScalars), hipHostMallocDefault));
        std::memset(snap_[i], 0, o.max_slots * sizeof(KScalars));
        TDDA_HIP(hipEventCreateWithFlags(&chunk_done_[i], hipEventDisableTiming));
        TDDA_HIP(hipEventCreateWithFlags(&rho_done_[i], hipEventDisableTiming));
    }
    TDDA_HIP(hipStreamSynchronize(main_));
    slots_.resize(o.max_slots);
    th_ = std::thread([this] { loop(); });
    adm_th_ = std::thread([this] { admit_loop(); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    cv_adm_.notify_all();
    if (adm_th_.joinable()) adm_th_.join();
    if (th_.joinable()) th_.join();
    (void)hipStreamSynchronize(main_);
    (void)hipStreamSynchronize(setup_);
    drop_graph(0);
    drop_graph(1);
    for (auto& rt : retiring_) {
        if (rt.ev) (void)hipEventDestroy(rt.ev);
        if (rt.host) (void)hipHostFree(rt.host);
    }
    retiring_.clear();
    for (auto& kv : pinned_fields_)
        for (double2* h : kv.second) (void)hipHostFree(h);
    for (auto& jr : joining_) (void)hipEventDestroy(jr.ev);
    joining_.clear();
    for (hipEvent_t e : free_events_) (void)hipEventDestroy(e);
    slots_.clear();
    pool_.clear();
    for (int i = 0; i < 2; ++i) {
        if (snap_[i]) (void)hipHostFree(snap_[i]);
        if (chunk_done_[i]) (void)hipEventDestroy(chunk_done_[i]);
        if (rho_done_[i]) (void)hipEventDestroy(rho_done_[i]);
    }
    (void)hipStreamDestroy(main_);
    (void)hipStreamDestroy(setup_);
}

double Server::now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

void Server::submit(ServeRequest&& r) {
    const long n = (long)r.desc.nx * r.desc.ny * r.desc.nz;
    TDDA_REQUIRE(n >= 1 && (long)r.rho.size() == n, ERR_INVALID, "density must hold nx*ny*nz values");
    r.t_submit = now();
    {
        std::lock_guard<std::mutex> lk(mu_);
        TDDA_REQUIRE(error_.empty(), ERR_STATE, "server stopped after an error: " + error_);
        pending_.push_back(std::move(r));
        stats_.submitted++;
    }
    cv_adm_.notify_all();
}

int Server::poll(std::vector<ServeResult>& out, int max) {
    std::lock_guard<std::mutex> lk(mu_);
    int k = 0;
    while (!done_.empty() && k < max) {
        out.push_back(done_.front());
        done_.pop_front();
        ++k;
    }
    return k;
}

int Server::take_field(long id, size_t n, std::vector<double2>& E) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = fields_.find(id);
    if (it == fields_.end()) return 1;
    if (it->second.size() != n) return 2;
    E = std::move(it->second);
    fields_.erase(it);
    return 0;
}

bool Server::idle() {
    std::lock_guard<std::mutex> lk(mu_);
    return pending_.empty() && busy_ == 0;
}

ServerStats Server::stats() {
    std::lock_guard<std::mutex> lk(mu_);
    return stats_;
}

std::string Server::error() {
    std::lock_guard<std::mutex> lk(mu_);
    return error_;
}

// Admit queued requests in arrival order while a slot is free and the memory
// estimate fits; their operators and densities are set up on the setup stream.
std::vector<int> Server::admit() {
    std::vector<int> joined;
    // rounds grow with the live count (8, 8, 16, 32, ...): the first chunks
    // start after a few set-ups instead of a full round of them
    int limit;
    {
        std::lock_guard<std::mutex> lk(mu_);
        limit = std::max(1, std::min(opt_.admit_per_round, std::max(8, (int)busy_ / 2)));
    }
    // this round's set-up buffers (their previous round has completed once its event has)
    const int r = rho_round_;
    if (rho_pending_[r]) {
        TDDA_HIP(hipEventSynchronize(rho_done_[r]));
        rho_pending_[r] = false;
    }
    GreenBatch& gb = green_batch_[r];
    try {
        admit_requests(limit, gb, joined);
    } catch (...) {
        gb.cancel();
        throw;
    }
    if (joined.empty()) return joined;
    rho_round_ ^= 1;
    // Green's set-ups of the round's new operators, then densities and
    // per-dipole set-up of every request, each in one upload and a few launches
    gb.flush(setup_);
    std::vector<Problem*> probs;
    std::vector<const double*> rho;
    for (int s : joined) {
        probs.push_back(slots_[s].prob.get());
        rho.push_back(slots_[s].req.rho.data());
    }
    round_setup_[r].run(probs, rho, setup_);
    for (int s : joined) {
        slots_[s].req.rho.clear();
        slots_[s].req.rho.shrink_to_fit();
    }
    TDDA_HIP(hipEventRecord(rho_done_[r], setup_));
    rho_pending_[r] = true;
    // hand the round to the loop thread: it joins once this event completes
    JoinRound jr;
    jr.slots = joined;
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (!free_events_.empty()) {
            jr.ev = free_events_.back();
            free_events_.pop_back();
        }
    }
    if (!jr.ev) TDDA_HIP(hipEventCreateWithFlags(&jr.ev, hipEventDisableTiming));
    TDDA_HIP(hipEventRecord(jr.ev, setup_));
    {
        std::lock_guard<std::mutex> lk(mu_);
        joining_.push_back(std::move(jr));
    }
    cv_.notify_all();
    return joined;
}

// Take queued requests into free slots (up to `limit`) and create their
// problems, queuing new operators' set-ups in gb.
void Server::admit_requests(int limit, GreenBatch& gb, std::vector<int>& joined) {
    while (true) {
        ServeRequest req;
        if ((int)joined.size() >= limit) break;
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (pending_.empty()) break;
            int free_slot = -1;
            for (int s = 0; s < (int)slots_.size(); ++s)
                if (!slots_[s].used) { free_slot = s; break; }
            if (free_slot < 0) break;
            const ProblemDesc& d = pending_.front().desc;
            const size_t need = Problem::estimate_bytes(d, opt_.refined, true);
            if (busy_ > 0 && device_bytes_in_use() + need > opt_.hbm_budget) {
                // drop cached operators nobody uses, then re-check
                OperatorCache::instance().trim(device_bytes_in_use() + need - opt_.hbm_budget);
                if (device_bytes_in_use() + need > opt_.hbm_budget) break;   // wait for memory
            }
            req = std::move(pending_.front());
            pending_.pop_front();
            busy_++;
            slots_[free_slot].used = true;
            slots_[free_slot].gen++;
            joined.push_back(free_slot);
        }
        // the slot is this thread's until its round is handed to the loop
        Slot& sl = slots_[joined.back()];
        const auto key = std::make_tuple(req.desc.nx, req.desc.ny, req.desc.nz);
        std::unique_ptr<System> reuse;
        {
            std::lock_guard<std::mutex> lk(pool_mu_);
            auto pit = pool_.find(key);
            if (pit != pool_.end() && !pit->second.empty()) {
                reuse = std::move(pit->second.back());
                pit->second.pop_back();
            }
        }
        sl.prob.reset(new Problem(req.desc, tables_, opt_.refined, setup_, nullptr, std::move(reuse), &gb));
        sl.prob->use_incident_r0();   // every solve starts from P = 0 with the incident field
        sl.prob->sys->set_jacobi(opt_.jacobi);
        sl.op_hit = sl.prob->op_hit ? 1 : 0;
        sl.t_admit = now();
        sl.req = std::move(req);
    }
}

// Admission thread: admit whenever requests wait and a slot is free; when
// memory holds admissions back, retry after a retirement (or 2 ms).
void Server::admit_loop() {
    try {
        while (true) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_adm_.wait_for(lk, std::chrono::milliseconds(20), [this] {
                    if (stop_ || !error_.empty()) return true;
                    if (pending_.empty()) return false;
                    for (const Slot& sl : slots_)
                        if (!sl.used) return true;
                    return false;
                });
                if (stop_ || !error_.empty()) break;
            }
            const double t0 = now();
            const std::vector<int> admitted = admit();
            {
                std::lock_guard<std::mutex> lk(mu_);
                stats_.host_ms[0] += 1e3 * (now() - t0);
            }
            if (admitted.empty()) {
                std::unique_lock<std::mutex> lk(mu_);
                if (stop_) break;
                cv_adm_.wait_for(lk, std::chrono::milliseconds(2));
            }
        }
    } catch (const std::exception& e) {
        fail(e.what());
    }
}

void Server::fail(const std::string& what) {
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (error_.empty()) error_ = what;
        busy_ = 0;
        pending_.clear();
    }
    cv_.notify_all();
    cv_adm_.notify_all();
}

// Finished requests leave their slot at once.  Without kept fields they are
// done; with them, the field is formed and copied to pinned memory on the
// compute stream behind an event, and the request completes when that copy
// has landed (complete_retirements), so the loop never waits for the stream.
void Server::retire(const std::vector<int>& fin, const KScalars* snap) {
    if (fin.empty()) return;
    const double t = now();
    std::vector<Retiring> out(fin.size());
    for (size_t k = 0; k < fin.size(); ++k) {
        Slot& sl = slots_[fin[k]];
        const KScalars& ks = snap[fin[k]];
        Retiring& rt = out[k];
        rt.r.id = sl.req.id;
        rt.r.status = ks.status;
        rt.r.iterations = ks.iterations;
        rt.r.n_ru = ks.n_ru;
        rt.r.err = ks.err;
        rt.r.err_est = ks.err_est;
        rt.r.t_submit = sl.req.t_submit;
        rt.r.t_admit = sl.t_admit;
        rt.r.op_hit = sl.op_hit;
        rt.shape = std::make_tuple(sl.req.desc.nx, sl.req.desc.ny, sl.req.desc.nz);
        rt.prob = std::move(sl.prob);
        // a preconditioned solve ends with Y = al P in P (the batch is gated
        // off for it, so later chunks leave it alone): the field is formed
        // from Y, and without kept fields nothing reads P again
        if (opt_.keep_fields) {
            rt.n = 3ull * rt.prob->ndip;
            rt.dev.alloc(rt.n);
            rt.prob->field(rt.dev.get(), main_, rt.prob->sys->jacobi());
            auto& pool = pinned_fields_[rt.n];
            if (!pool.empty()) {
                rt.host = pool.back();
                pool.pop_back();
            } else {
                TDDA_HIP(hipHostMalloc((void**)&rt.host, rt.n * sizeof(double2), hipHostMallocDefault));
            }
            TDDA_HIP(hipMemcpyAsync(rt.host, rt.dev.get(), rt.n * sizeof(double2), hipMemcpyDeviceToHost, main_));
            TDDA_HIP(hipEventCreateWithFlags(&rt.ev, hipEventDisableTiming));
            TDDA_HIP(hipEventRecord(rt.ev, main_));
        }
    }
    {
        std::lock_guard<std::mutex> lk(mu_);
        for (int f : fin) slots_[f].used = false;
    }
    cv_adm_.notify_all();
    for (auto& rt : out) {
        if (opt_.keep_fields) retiring_.push_back(std::move(rt));
        else finish(rt, t);
    }
}

void Server::finish(Retiring& rt, double t) {
    rt.r.t_done = t;
    std::vector<double2> field;
    if (rt.host) {
        field.assign(rt.host, rt.host + rt.n);
        pinned_fields_[rt.n].push_back(rt.host);
        rt.host = nullptr;
    }
    if (rt.ev) (void)hipEventDestroy(rt.ev);
    rt.ev = nullptr;
    rt.dev.release();
    rt.prob->sys->op.reset();   // let the operator cache evict it
    {
        std::lock_guard<std::mutex> lk(pool_mu_);
        pool_[rt.shape].push_back(std::move(rt.prob->sys));
    }
    rt.prob.reset();
    {
        std::lock_guard<std::mutex> lk(mu_);
        done_.push_back(rt.r);
        if (opt_.keep_fields) fields_[rt.r.id] = std::move(field);
        busy_--;
        stats_.completed++;
    }
    cv_adm_.notify_all();   // memory was freed
}

void Server::complete_retirements(bool wait) {
    while (!retiring_.empty()) {
        Retiring& rt = retiring_.front();
        if (wait) {
            TDDA_HIP(hipEventSynchronize(rt.ev));
        } else {
            const hipError_t e = hipEventQuery(rt.ev);
            if (e == hipErrorNotReady) break;
            TDDA_HIP(e);
        }
        finish(rt, now());
        retiring_.pop_front();
    }
}

// (the admission thread only appends rounds, this thread only removes them)
std::vector<int> Server::ready_joiners(bool wait) {
    std::vector<int> ready;
    while (true) {
        hipEvent_t ev = nullptr;
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (joining_.empty()) break;
            ev = joining_.front().ev;
        }
        if (wait && ready.empty()) {
            TDDA_HIP(hipEventSynchronize(ev));
        } else {
            const hipError_t e = hipEventQuery(ev);
            if (e == hipErrorNotReady) break;
            TDDA_HIP(e);
        }
        std::lock_guard<std::mutex> lk(mu_);
        JoinRound& jr = joining_.front();
        ready.insert(ready.end(), jr.slots.begin(), jr.slots.end());
        free_events_.push_back(jr.ev);
        joining_.pop_front();
    }
    return ready;
}

void Server::drop_graph(int b) {
    if (chunk_graph_[b]) (void)hipGraphExecDestroy(chunk_graph_[b]);
    chunk_graph_[b] = nullptr;
}

// A chunk of iterations of batch b and the end-of-chunk update, on main_.
bool Server::enqueue_chunk(int b) {
    const TwiddlePools tw = PlanRegistry::instance().pools();
    const LaunchSet& ls = batch_[b].ls;
    auto body = [&](hipStream_t st) {
        for (int k = 0; k < opt_.chunk; ++k) launch_iteration(ls, tw, opt_.refined, st);
        launch_finalize(ls, st);
    };
    batch_uses_[b]++;
    if (graph_capture() && batch_uses_[b] >= 4) {   // recording pays back over a few replays
        if (!chunk_graph_[b]) chunk_graph_[b] = capture_graph(body);
        TDDA_HIP(hipGraphLaunch(chunk_graph_[b], main_));
        return true;
    }
    body(main_);
    return false;
}

void Server::loop() {
    try {
        int cur = 0;   // batch whose chunk is enqueued next
        while (true) {
            // 1. enqueue a chunk of iterations for batch `cur`
            if (!members_[cur].empty()) {
                const double te = now();
                const bool replay = enqueue_chunk(cur);
                TDDA_HIP(hipMemcpyAsync(snap_[cur], ks_.get(), opt_.max_slots * sizeof(KScalars),
                                        hipMemcpyDeviceToHost, main_));
                TDDA_HIP(hipEventRecord(chunk_done_[cur], main_));
                inflight_[cur] = true;
                std::lock_guard<std::mutex> lk(mu_);
                stats_.enqueue_ms += 1e3 * (now() - te);
                if (replay) stats_.graph_chunks++;
                stats_.chunks++;
                stats_.launches += (long)opt_.chunk * (opt_.refined ? 18 : 12) + 3;
            }
            // 2. statuses of the previous chunk (the current one keeps running)
            double tm[4];
            tm[0] = now();
            const int prev = 1 - cur;
            std::vector<int> fin;
            if (inflight_[prev]) {
                TDDA_HIP(hipEventSynchronize(chunk_done_[prev]));
                inflight_[prev] = false;
                // (a slot retired after an earlier chunk may already hold a
                // new request: only the generation that ran in that chunk counts)
                std::lock_guard<std::mutex> lk(mu_);
                for (const auto& m : members_[prev])
                    if (slots_[m.first].used && slots_[m.first].gen == m.second &&
                        snap_[prev][m.first].status != JOB_ACTIVE)
                        fin.push_back(m.first);
            }
            tm[1] = now();
            {
                std::lock_guard<std::mutex> lk(mu_);
                stats_.peak_bytes = std::max(stats_.peak_bytes, device_bytes_in_use());
            }
            // jobs still in the running batch are gated off once finished,
            // so their slots can be recycled now (their buffers once their
            // field has been copied out)
            retire(fin, snap_[prev]);
            complete_retirements(false);
            tm[2] = now();
            // 3. the other batch for the next chunk: every live request (the
            // live members of the current batch, which holds all joined
            // before it was built) and the admitted requests whose set-up has
            // completed (with nothing else to run, the oldest round is waited for)
            std::vector<std::pair<int, unsigned>> next;
            size_t dead = 0;
            {
                std::lock_guard<std::mutex> lk(mu_);
                for (const auto& m : members_[cur])
                    if (slots_[m.first].used && slots_[m.first].gen == m.second) next.push_back(m);
                for (const auto& m : members_[prev])
                    if (!slots_[m.first].used || slots_[m.first].gen != m.second) ++dead;
            }
            const std::vector<int> joined = ready_joiners(next.empty());
            if (!joined.empty()) {
                live_version_++;
                std::vector<System*> js;
                std::vector<KScalars*> jk;
                for (int s : joined) {
                    js.push_back(slots_[s].prob->sys.get());
                    jk.push_back(ks_.get() + s);
                }
                KScalars in;
                std::memset(&in, 0, sizeof(KScalars));
                in.tol = opt_.tol;
                in.maxit = opt_.maxit;
                in.refined = opt_.refined ? 1 : 0;
                in.ru_delta = opt_.ru_delta;
                in.tol_e = opt_.refined ? opt_.err_tol : 0.0;
                in.status = JOB_ACTIVE;
                init_batch_[prev].build(js, jk, opt_.refined, main_, false);
                launch_ks_template(init_batch_[prev].ls, in, main_);
                launch_init(init_batch_[prev].ls, false, main_);
                std::lock_guard<std::mutex> lk(mu_);
                for (int s : joined) next.emplace_back(s, slots_[s].gen);
            }
            const bool stale = batch_version_[prev] != live_version_ || members_[prev].empty() ||
                               2 * dead > members_[prev].size();
            if (stale) {
                members_[prev] = next;
                drop_graph(prev);
                batch_uses_[prev] = 0;
                batch_version_[prev] = live_version_;
                if (!next.empty()) {
                    std::vector<System*> sys;
                    std::vector<KScalars*> ks;
                    for (const auto& m : next) {
                        sys.push_back(slots_[m.first].prob->sys.get());
                        ks.push_back(ks_.get() + m.first);
                    }
                    batch_[prev].build(sys, ks, opt_.refined, main_);
                    std::lock_guard<std::mutex> lk(mu_);
                    stats_.rebuilds++;
                    stats_.max_active = std::max(stats_.max_active, (int)next.size());
                }
            }
            if (next.empty()) members_[prev].clear();   // nothing live: the loop idles
            tm[3] = now();
            {
                std::lock_guard<std::mutex> lk(mu_);
                for (int q = 0; q < 3; ++q) stats_.host_ms[q + 1] += 1e3 * (tm[q + 1] - tm[q]);
            }
            cur = prev;
            // 4. idle: nothing running, in flight or being set up
            std::unique_lock<std::mutex> lk(mu_);
            if (stop_ || !error_.empty()) break;
            if (members_[cur].empty() && !inflight_[0] && !inflight_[1] && joining_.empty()) {
                lk.unlock();
                complete_retirements(true);
                lk.lock();
                cv_.wait_for(lk, std::chrono::milliseconds(50),
                             [this] { return stop_ || !error_.empty() || !joining_.empty(); });
                if (stop_ || !error_.empty()) break;
            }
        }
        (void)hipStreamSynchronize(main_);
        complete_retirements(true);
    } catch (const std::exception& e) {
        fail(e.what());
    }
}

}  // namespace tdda
