// Host runtime implementation: memory, plans, operators, systems, batches.
#include "runtime.h"
#include "fft_fast.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <tuple>

namespace tdda {

// ---------------------------------------------------------------------------
// memory
// ---------------------------------------------------------------------------
// Caching allocator: hipFree synchronises the device, and the server creates
// and retires systems continuously, so freed blocks go to per-size free lists
// and are reused.  Callers free a block only after the work using it is done.
namespace {
std::mutex g_mem_mu;
std::map<size_t, std::vector<void*>> g_free;
size_t g_in_use = 0, g_cached = 0;

size_t round_size(size_t b) {
    if (b >= (1u << 20)) return (b + (2u << 20) - 1) & ~size_t((2u << 20) - 1);
    return (b + 511) & ~size_t(511);
}

void release_cached_locked() {
    for (auto& kv : g_free)
        for (void* p : kv.second) (void)hipFree(p);
    g_free.clear();
    g_cached = 0;
}
}  // namespace

size_t device_bytes_in_use() {
    std::lock_guard<std::mutex> lk(g_mem_mu);
    return g_in_use;
}

size_t device_bytes_cached() {
    std::lock_guard<std::mutex> lk(g_mem_mu);
    return g_cached;
}

void device_release_cached() {
    std::lock_guard<std::mutex> lk(g_mem_mu);
    release_cached_locked();
}

void* dev_alloc(size_t bytes) {
    const size_t sz = round_size(bytes);
    std::lock_guard<std::mutex> lk(g_mem_mu);
    auto it = g_free.find(sz);
    if (it != g_free.end() && !it->second.empty()) {
        void* p = it->second.back();
        it->second.pop_back();
        g_cached -= sz;
        g_in_use += sz;
        return p;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, sz);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        release_cached_locked();
        e = hipMalloc(&p, sz);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error(ERR_MEMORY, "hipMalloc of " + std::to_string(sz) + " bytes failed: " + hipGetErrorString(e));
    }
    g_in_use += sz;
    return p;
}

void dev_free(void* p, size_t bytes) {
    if (!p) return;
    const size_t sz = round_size(bytes);
    std::lock_guard<std::mutex> lk(g_mem_mu);
    g_free[sz].push_back(p);
    g_in_use -= sz;
    g_cached += sz;
}

hipStream_t default_stream() { return nullptr; }

// ---------------------------------------------------------------------------
// FFT lengths and plans
// ---------------------------------------------------------------------------
static bool smooth235(int L) {
    for (int p : {2, 3, 5})
        while (L % p == 0) L /= p;
    return L == 1;
}

int choose_fft_length(int n) {
    int L = std::max(2, 2 * n - 1);
    while (!smooth235(L)) ++L;
    return L;
}

std::vector<int> radix_sequence(int L) {
    std::vector<int> r;
    int a = 0;
    while (L % 2 == 0) { L /= 2; ++a; }
    while (a >= 3) { r.push_back(8); a -= 3; }
    if (a == 2) r.push_back(4);
    if (a == 1) r.push_back(2);
    while (L % 5 == 0) { r.push_back(5); L /= 5; }
    while (L % 3 == 0) { r.push_back(3); L /= 3; }
    if (L != 1) throw Error(ERR_INVALID, "FFT length is not 2,3,5-smooth");
    return r;
}

PlanRegistry& PlanRegistry::instance() {
    static PlanRegistry r;
    return r;
}

static constexpr int kTwCapacity = 1 << 18;

FftPlan PlanRegistry::plan(int L) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = plans_.find(L);
    if (it != plans_.end()) return it->second;
    if (!tw64_.get()) {
        tw64_.alloc(kTwCapacity);
        tw32_.alloc(kTwCapacity);
    }
    TDDA_REQUIRE(used_ + L <= kTwCapacity, ERR_MEMORY, "twiddle pool exhausted");
    std::vector<int> rs = radix_sequence(L);
    TDDA_REQUIRE((int)rs.size() <= FFT_MAX_STAGES, ERR_INVALID, "too many FFT stages");
    FftPlan p{};
    p.L = L;
    p.nst = (int)rs.size();
    for (int i = 0; i < p.nst; ++i) p.radix[i] = rs[i];
    p.tw_off = used_;
    std::vector<double2> h64(L);
    std::vector<float2> h32(L);
    for (int m = 0; m < L; ++m) {
        const long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)m / (long double)L;
        h64[m] = make_double2((double)cosl(ang), (double)sinl(ang));
        h32[m] = make_float2((float)h64[m].x, (float)h64[m].y);
    }
    TDDA_HIP(hipMemcpy(tw64_.get() + used_, h64.data(), L * sizeof(double2), hipMemcpyHostToDevice));
    TDDA_HIP(hipMemcpy(tw32_.get() + used_, h32.data(), L * sizeof(float2), hipMemcpyHostToDevice));
    used_ += L;
    plans_[L] = p;
    return p;
}

TwiddlePools PlanRegistry::pools() {
    std::lock_guard<std::mutex> lk(mu_);
    TwiddlePools t;
    t.tw32 = tw32_.get();
    t.tw64 = tw64_.get();
    return t;
}

// ---------------------------------------------------------------------------
// Si/Ci tables
// ---------------------------------------------------------------------------
namespace {
struct SiciEntry {
    DevBuf<double> si, ci;
    long n;
    double delta;
};
std::mutex g_sici_mu;
std::map<uint64_t, std::unique_ptr<SiciEntry>> g_sici;

uint64_t fnv64(const void* data, size_t bytes, uint64_t h) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < bytes; ++i) { h ^= p[i]; h *= 1099511628211ull; }
    return h;
}
}  // namespace

SiciDevice sici_tables(const double* si, const double* ci, long n, double delta) {
    SiciDevice out;
    out.n = n;
    out.delta = delta;
    if (n <= 0 || si == nullptr || ci == nullptr) return out;   // LDR needs none
    uint64_t h = 1469598103934665603ull;
    h = fnv64(&n, sizeof n, h);
    h = fnv64(&delta, sizeof delta, h);
    h = fnv64(si, n * sizeof(double), h);
    h = fnv64(ci, n * sizeof(double), h);
    std::lock_guard<std::mutex> lk(g_sici_mu);
    auto it = g_sici.find(h);
    if (it == g_sici.end()) {
        auto e = std::make_unique<SiciEntry>();
        e->si.alloc(n);
        e->ci.alloc(n);
        TDDA_HIP(hipMemcpy(e->si.get(), si, n * sizeof(double), hipMemcpyHostToDevice));
        TDDA_HIP(hipMemcpy(e->ci.get(), ci, n * sizeof(double), hipMemcpyHostToDevice));
        e->n = n;
        e->delta = delta;
        it = g_sici.emplace(h, std::move(e)).first;
    }
    out.si = it->second->si.get();
    out.ci = it->second->ci.get();
    out.id = h;
    return out;
}

// ---------------------------------------------------------------------------
// operators
// ---------------------------------------------------------------------------
namespace {
// Timing events of operator set-ups, reused: the server creates and evicts
// operators continuously, and event creation costs host time per request.
std::mutex g_ev_mu;
std::vector<hipEvent_t> g_ev_free;

hipEvent_t take_event() {
    {
        std::lock_guard<std::mutex> lk(g_ev_mu);
        if (!g_ev_free.empty()) {
            hipEvent_t e = g_ev_free.back();
            g_ev_free.pop_back();
            return e;
        }
    }
    hipEvent_t e = nullptr;
    TDDA_HIP(hipEventCreate(&e));
    return e;
}

void give_event(hipEvent_t e) {
    std::lock_guard<std::mutex> lk(g_ev_mu);
    g_ev_free.push_back(e);
}
}  // namespace

bool OperatorKey::operator<(const OperatorKey& o) const {
    return std::tie(nx, ny, nz, d, K, method, maxm, maxn, lm, ln, sici_id) <
           std::tie(o.nx, o.ny, o.nz, o.d, o.K, o.method, o.maxm, o.maxn, o.lm, o.ln, o.sici_id);
}

Operator::Operator(const OperatorKey& k, const SiciDevice& tables, hipStream_t st, GreenBatch* batch) : key(k) {
    nx = k.nx; ny = k.ny; nz = k.nz;
    TDDA_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, ERR_INVALID, "grid dimensions must be positive");
    lx = choose_fft_length(nx);
    ly = choose_fft_length(ny);
    lz = choose_fft_length(nz);
    TDDA_REQUIRE(lx <= MAX_LXY && ly <= MAX_LXY && lz <= MAX_LZ, ERR_INVALID,
                 "grid too large for one device transform (Nx, Ny, Nz <= 128)");
    hx = lx / 2 + 1; hy = ly / 2 + 1; hz = lz / 2 + 1;
    PlanRegistry& reg = PlanRegistry::instance();
    px = reg.plan(lx);
    py = reg.plan(ly);
    pz = reg.plan(lz);
    TwiddlePools tw = reg.pools();
    if (k.method == 0)
        TDDA_REQUIRE(tables.si != nullptr && tables.n > 0, ERR_INVALID, "FCD needs Si/Ci tables");

    GreenParams& P = gp_;
    P.nx = nx; P.ny = ny; P.nz = nz;
    P.d = k.d; P.K = k.K; P.method = k.method;
    P.maxm = k.maxm; P.maxn = k.maxn; P.lm = k.lm; P.ln = k.ln;
    P.nsici = tables.n; P.delta = tables.delta;
    si_ = tables.si;
    ci_ = tables.ci;

    const long noct = (long)nx * ny * nz;
    tmp_.resize(4);
    tmp_[0].alloc(6 * noct);
    tmp_[1].alloc(6L * nz * ny * hx);
    tmp_[2].alloc(6L * nz * hy * hx);
    tmp_[3].alloc(6L * hz * hy * hx);
    green64.alloc(6L * hz * hy * hx);
    green32.alloc(6L * hz * hy * hx);
    if (batch) {   // the batch runs the other steps (its flush records done_)
        if (P.maxm != 0 || P.maxn != 0) launch_green_octant(P, si_, ci_, tmp_[0].get(), st);
        return;
    }
    start_ = take_event();
    done_ = take_event();
    TDDA_HIP(hipEventRecord(start_, st));
    launch_green_octant(P, si_, ci_, tmp_[0].get(), st);
    launch_green_dft(tmp_[0].get(), tmp_[1].get(), nz * ny, nx, hx, 1, lx, 0, tw.tw64 + px.tw_off, st);
    launch_green_dft(tmp_[1].get(), tmp_[2].get(), nz, ny, hy, hx, ly, 1, tw.tw64 + py.tw_off, st);
    launch_green_dft(tmp_[2].get(), tmp_[3].get(), 1, nz, hz, hy * hx, lz, 2, tw.tw64 + pz.tw_off, st);
    launch_green_interleave(tmp_[3].get(), (long)hz * hy, hx, green64.get(), green32.get(), st);
    TDDA_HIP(hipEventRecord(done_, st));
}

Operator::~Operator() {
    // the temporaries and the spectrum may still be in use by the build
    if (done_) (void)hipEventSynchronize(done_);
    tmp_.clear();
    if (start_) give_event(start_);
    if (done_) give_event(done_);
}

double Operator::setup_ms() const {
    if (!start_) return -1.0;
    TDDA_HIP(hipEventSynchronize(done_));
    float ms = 0.f;
    TDDA_HIP(hipEventElapsedTime(&ms, start_, done_));
    return ms;
}

// ---------------------------------------------------------------------------
// batched Green's set-ups
// ---------------------------------------------------------------------------
GreenBatch::~GreenBatch() {
    cancel();
    if (flushed_) {
        (void)hipEventSynchronize(flushed_);
        (void)hipEventDestroy(flushed_);
    }
}

void GreenBatch::add(const std::shared_ptr<Operator>& op, hipStream_t st) {
    ops_.push_back(op);
    st_ = st;
}

std::shared_ptr<Operator> GreenBatch::find(const OperatorKey& key) const {
    for (const auto& op : ops_)
        if (!(op->key < key) && !(key < op->key)) return op;
    return nullptr;
}

void GreenBatch::cancel() {
    if (ops_.empty()) return;
    if (st_) (void)hipStreamSynchronize(st_);   // queued lattice sums write the temporaries
    OperatorCache::instance().publish(ops_, false);
    ops_.clear();
}

void GreenBatch::flush(hipStream_t st) {
    if (ops_.empty()) return;
    const TwiddlePools tw = PlanRegistry::instance().pools();
    std::vector<GreenSingleJob> single;
    std::vector<GreenDftJob> dft[3][2];   // axis, 8 or 16 outputs per thread
    std::vector<GreenInterleaveJob> inter;
    int bs = 0, bi = 0, bd[3][2] = {};
    for (const auto& sp : ops_) {
        Operator& o = *sp;
        if (o.gp_.maxm == 0 && o.gp_.maxn == 0) {
            single.push_back(GreenSingleJob{o.gp_, o.si_, o.ci_, o.tmp_[0].get(), bs});
            bs += green_single_blocks(o.gp_);
        }
        // the three axes of the direct build, in its order
        const struct {
            const double2* in;
            double2* out;
            int outer, no, h, inner, L;
            const double2* tw;
        } ax[3] = {{o.tmp_[0].get(), o.tmp_[1].get(), o.nz * o.ny, o.nx, o.hx, 1, o.lx, tw.tw64 + o.px.tw_off},
                   {o.tmp_[1].get(), o.tmp_[2].get(), o.nz, o.ny, o.hy, o.hx, o.ly, tw.tw64 + o.py.tw_off},
                   {o.tmp_[2].get(), o.tmp_[3].get(), 1, o.nz, o.hz, o.hy * o.hx, o.lz, tw.tw64 + o.pz.tw_off}};
        for (int a = 0; a < 3; ++a) {
            const int kb = green_dft_kb(ax[a].h), q = kb == 8 ? 0 : 1;
            dft[a][q].push_back(GreenDftJob{ax[a].in, ax[a].out, ax[a].tw, ax[a].outer, ax[a].no, ax[a].h,
                                            ax[a].inner, ax[a].L, a, (ax[a].h + kb - 1) / kb, bd[a][q]});
            bd[a][q] += green_dft_blocks(ax[a].outer, ax[a].h, ax[a].inner);
        }
        inter.push_back(GreenInterleaveJob{o.tmp_[3].get(), o.green64.get(), o.green32.get(), (long)o.hz * o.hy,
                                           o.hx, bi});
        bi += green_interleave_blocks((long)o.hz * o.hy, o.hx);
    }
    // the job tables back to back (256-byte aligned) in one upload
    auto al256 = [](size_t b) { return (b + 255) & ~size_t(255); };
    size_t bytes = al256(single.size() * sizeof(GreenSingleJob)) + al256(inter.size() * sizeof(GreenInterleaveJob));
    for (auto& a : dft)
        for (auto& v : a) bytes += al256(v.size() * sizeof(GreenDftJob));
    if (flushed_) TDDA_HIP(hipEventSynchronize(flushed_));   // the tables below are reused
    char* h = stage_.reserve(bytes);
    if (dev_.size() < bytes) dev_.alloc(std::max(bytes, 2 * dev_.size()));
    size_t off = 0;
    auto put = [&](const void* src, size_t n) {
        if (n) std::memcpy(h + off, src, n);
        const size_t at = off;
        off += al256(n);
        return dev_.get() + at;
    };
    const auto* d_single = reinterpret_cast<const GreenSingleJob*>(put(single.data(), single.size() * sizeof(GreenSingleJob)));
    const GreenDftJob* d_dft[3][2];
    for (int a = 0; a < 3; ++a)
        for (int q = 0; q < 2; ++q)
            d_dft[a][q] = reinterpret_cast<const GreenDftJob*>(put(dft[a][q].data(), dft[a][q].size() * sizeof(GreenDftJob)));
    const auto* d_inter = reinterpret_cast<const GreenInterleaveJob*>(put(inter.data(), inter.size() * sizeof(GreenInterleaveJob)));
    TDDA_HIP(hipMemcpyAsync(dev_.get(), h, off, hipMemcpyHostToDevice, st));
    launch_green_single_batch(d_single, (int)single.size(), bs, st);
    for (int a = 0; a < 3; ++a)
        for (int q = 0; q < 2; ++q) launch_green_dft_batch(d_dft[a][q], (int)dft[a][q].size(), bd[a][q], q ? 16 : 8, st);
    launch_green_interleave_batch(d_inter, (int)inter.size(), bi, st);
    if (!flushed_) TDDA_HIP(hipEventCreateWithFlags(&flushed_, hipEventDisableTiming));
    TDDA_HIP(hipEventRecord(flushed_, st));
    for (const auto& sp : ops_) {
        sp->done_ = take_event();
        TDDA_HIP(hipEventRecord(sp->done_, st));
    }
    OperatorCache::instance().publish(ops_, true);
    ops_.clear();
}

bool Operator::reap() {
    if (tmp_.empty()) return true;
    const hipError_t e = hipEventQuery(done_);
    if (e == hipErrorNotReady) return false;
    TDDA_HIP(e);
    tmp_.clear();
    return true;
}

OperatorCache& OperatorCache::instance() {
    static OperatorCache c;
    return c;
}

size_t OperatorCache::bytes() const {
    std::lock_guard<std::mutex> lk(mu_);
    size_t s = 0;
    for (auto& kv : map_) s += kv.second->bytes();
    return s;
}

size_t OperatorCache::trim(size_t bytes) {
    std::lock_guard<std::mutex> lk(mu_);
    size_t freed = 0;
    for (size_t i = 0; i < lru_.size() && freed < bytes;) {
        auto it = map_.find(lru_[i]);
        if (it != map_.end() && it->second.use_count() == 1) {
            freed += it->second->bytes();
            map_.erase(it);
            lru_.erase(lru_.begin() + i);
        } else {
            ++i;
        }
    }
    return freed;
}

void OperatorCache::clear() {
    std::lock_guard<std::mutex> lk(mu_);
    map_.clear();
    lru_.clear();
}

// Builds finish in the order they were enqueued (mostly on one stream), so
// the scan stops at the first build still running: a cache lookup then costs
// one event query, not one per operator of the last admission rounds.
void OperatorCache::reap_locked() {
    size_t i = 0;
    for (; i < unreaped_.size(); ++i) {
        std::shared_ptr<Operator> op = unreaped_[i].lock();
        if (op && !op->reap()) break;
    }
    unreaped_.erase(unreaped_.begin(), unreaped_.begin() + i);
}

void OperatorCache::touch_locked(const OperatorKey& key) {
    auto pos = std::find_if(lru_.begin(), lru_.end(),
                            [&](const OperatorKey& k) { return !(k < key) && !(key < k); });
    if (pos != lru_.end()) lru_.erase(pos);
    lru_.push_back(key);
}

// One thread builds a missing operator; others asking for the same key wait
// for it instead of building a second copy.
std::shared_ptr<Operator> OperatorCache::get(const OperatorKey& key, const SiciDevice& tables, hipStream_t st,
                                             bool* hit, GreenBatch* batch) {
    if (batch) {   // queued earlier in the same batch: same stream, set up by the same flush
        std::shared_ptr<Operator> q = batch->find(key);
        if (q) {
            if (hit) *hit = true;
            return q;
        }
    }
    {
        std::unique_lock<std::mutex> lk(mu_);
        if (batch && !batch->empty() && building_.count(key)) {
            // another thread builds this key: publish this batch's queued
            // builds before waiting, so that two batching threads never wait
            // for each other's keys
            lk.unlock();
            batch->flush(st);
            lk.lock();
        }
        built_.wait(lk, [&] { return building_.count(key) == 0; });
        reap_locked();
        auto it = map_.find(key);
        if (it != map_.end()) {
            touch_locked(key);
            if (hit) *hit = true;
            // a build enqueued on another stream may still run
            TDDA_HIP(hipStreamWaitEvent(st, it->second->ready(), 0));
            return it->second;
        }
        building_.insert(key);
    }
    if (hit) *hit = false;
    std::shared_ptr<Operator> op;
    try {
        op = std::make_shared<Operator>(key, tables, st, batch);
    } catch (...) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            building_.erase(key);
        }
        built_.notify_all();
        throw;
    }
    if (batch) {   // stays "being built" until the batch's flush publishes it
        batch->add(op, st);
        return op;
    }
    {
        std::lock_guard<std::mutex> lk(mu_);
        building_.erase(key);
        insert_locked(key, op);
    }
    built_.notify_all();
    return op;
}

void OperatorCache::insert_locked(const OperatorKey& key, const std::shared_ptr<Operator>& op) {
    map_[key] = op;
    unreaped_.push_back(op);
    touch_locked(key);
    size_t total = 0;
    for (auto& kv : map_) total += kv.second->bytes();
    // evict least recently used entries nobody else holds
    for (size_t i = 0; i < lru_.size() && total > budget_;) {
        auto it = map_.find(lru_[i]);
        if (it != map_.end() && it->second.use_count() == 1 && it->second != op) {
            total -= it->second->bytes();
            map_.erase(it);
            lru_.erase(lru_.begin() + i);
        } else {
            ++i;
        }
    }
}

void OperatorCache::publish(const std::vector<std::shared_ptr<Operator>>& ops, bool built) {
    {
        std::lock_guard<std::mutex> lk(mu_);
        for (const auto& op : ops) {
            building_.erase(op->key);
            if (built) insert_locked(op->key, op);
        }
    }
    built_.notify_all();
}

// ---------------------------------------------------------------------------
// systems
// ---------------------------------------------------------------------------
static size_t grid_elems(const Operator& op) { return (size_t)3 * op.nz * op.ly * op.lx; }

size_t System::estimate_bytes(const Operator& op, int ndip, bool refined) {
    const size_t n3 = 3ull * ndip;
    const size_t ks = refined ? sizeof(float2) : sizeof(double2);
    size_t b = grid_elems(op) * sizeof(double2);
    b += 6 * n3 * ks;
    b += 3 * n3 * sizeof(double2);   // P, rhs, scratch64
    b += 2 * ndip * (sizeof(double2) + sizeof(float2));
    b += (size_t)std::max<long>(n3 / (FFT_THREADS * UPD_EPT) + 1, (long)op.ny * op.nz) * 6 * sizeof(double);
    if (refined) b += (size_t)op.ny * op.nz * XF_NACC * sizeof(double) + ndip * sizeof(float2);
    return b;
}

System::System(std::shared_ptr<Operator> op_, int ndip_, const int* box_map_host, bool refined_, hipStream_t st)
    : op(std::move(op_)), ndip(ndip_), refined(refined_) {
    const Operator& o = *op;
    TDDA_REQUIRE(ndip >= 1, ERR_INVALID, "system needs at least one dipole");
    const size_t n3 = 3ull * ndip;
    const size_t kbytes = refined ? sizeof(float2) : sizeof(double2);
    grid.alloc(grid_elems(o) * sizeof(double2));
    kvec.alloc((refined ? 7 : 6) * n3 * kbytes);
    P.alloc(n3);
    rhs.alloc(n3);
    scratch64.alloc(n3);
    al64.alloc(ndip);
    al32.alloc(ndip);
    ial64.alloc(ndip);
    ial32.alloc(ndip);
    P.zero(st);

    std::memset(&job, 0, sizeof job);
    job.nx = o.nx; job.ny = o.ny; job.nz = o.nz; job.ndip = ndip;
    job.lx = o.lx; job.ly = o.ly; job.lz = o.lz;
    job.hx = o.hx; job.hy = o.hy; job.hz = o.hz;
    job.px = o.px; job.py = o.py; job.pz = o.pz;
    plan_tiles(job);
    // (c64 plane passes write one X inverse partial per z-plane)
    const long nparts = std::max({(long)job.tupd, (long)job.tx32, (long)job.tx64, (long)job.nz});
    // reliable updates run the FP64 X forward pass: its tiles hold the
    // field-error partials
    const long nxf = refined ? (long)job.txf64 * XF_NACC : 0;
    partial.alloc(nparts * PARTIAL_STRIDE + nxf);
    if (refined) efac.alloc(ndip);
    if (box_map_host) {
        box_map.alloc((size_t)o.nx * o.ny * o.nz);
        box_map.upload(box_map_host, box_map.size(), st);
        job.box_map = box_map.get();
    } else {
        TDDA_REQUIRE((long)ndip == (long)o.nx * o.ny * o.nz, ERR_INVALID,
                     "identity dipole map needs the full box");
        job.box_map = nullptr;
    }
    job.green32 = o.green32.get();
    job.green64 = o.green64.get();
    job.al32 = al32.get();
    job.al64 = al64.get();
    job.ial32 = ial32.get();
    job.ial64 = ial64.get();
    job.jacobi = 0;
    job.grid = grid.get();
    char* k = kvec.get();
    job.r = k;
    job.r0 = k + 1 * n3 * kbytes;
    job.p = k + 2 * n3 * kbytes;
    job.v = k + 3 * n3 * kbytes;
    job.t = k + 4 * n3 * kbytes;
    job.s = k + 5 * n3 * kbytes;
    job.dP = refined ? static_cast<void*>(k + 6 * n3 * kbytes) : nullptr;
    job.P = P.get();
    job.rhs = rhs.get();
    job.partial = partial.get();
    job.partial_xf = refined ? partial.get() + nparts * PARTIAL_STRIDE : nullptr;
    job.efac = nullptr;   // f = 1 until set_field_factor
    job.ks = nullptr;
}

size_t System::bytes() const {
    return grid.bytes() + kvec.bytes() + P.bytes() + rhs.bytes() + scratch64.bytes() + al64.bytes() +
           al32.bytes() + ial64.bytes() + ial32.bytes() + partial.bytes() + efac.bytes() + box_map.bytes();
}

namespace {
// c64 copy of 1/alpha and alpha itself in both precisions
__global__ void k_alpha_derived(const double2* __restrict__ a, float2* __restrict__ a32, double2* __restrict__ ia64,
                                float2* __restrict__ ia32, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 v = a[i];
    a32[i] = make_float2((float)v.x, (float)v.y);
    const double2 w = ddiv(make_double2(1.0, 0.0), v);
    ia64[i] = w;
    ia32[i] = make_float2((float)w.x, (float)w.y);
}

__global__ void k_scale_solution(double2* __restrict__ P, const double2* __restrict__ f, long n3) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n3) P[i] = dcmul(P[i], f[i / 3]);
}
}  // namespace

void System::set_alpha(const double2* al_dip_host, hipStream_t st) {
    al64.upload(al_dip_host, ndip, st);
    hipLaunchKernelGGL(k_alpha_derived, dim3((ndip + 255) / 256), dim3(256), 0, st, al64.get(), al32.get(),
                       ial64.get(), ial32.get(), (long)ndip);
    TDDA_HIP(hipGetLastError());
}

void System::set_alpha_device(const double2* al_dip_dev, hipStream_t st) {
    TDDA_HIP(hipMemcpyAsync(al64.get(), al_dip_dev, ndip * sizeof(double2), hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_alpha_derived, dim3((ndip + 255) / 256), dim3(256), 0, st, al64.get(), al32.get(),
                       ial64.get(), ial32.get(), (long)ndip);
    TDDA_HIP(hipGetLastError());
}

void System::scale_solution(bool to_y, hipStream_t st) {
    const long n3 = 3L * ndip;
    hipLaunchKernelGGL(k_scale_solution, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, st, P.get(),
                       to_y ? al64.get() : ial64.get(), n3);
    TDDA_HIP(hipGetLastError());
}

void System::set_field_factor(const double2* f, hipStream_t st) {
    if (!refined) return;
    std::vector<float2> h(ndip);
    for (int i = 0; i < ndip; ++i) h[i] = make_float2((float)f[i].x, (float)f[i].y);
    efac.upload(h.data(), ndip, st);
    TDDA_HIP(hipStreamSynchronize(st));   // h leaves scope
    job.efac = efac.get();
}

void System::set_rhs(const double2* h, hipStream_t st) { rhs.upload(h, 3ull * ndip, st); }
void System::set_solution(const double2* h, hipStream_t st) { P.upload(h, 3ull * ndip, st); }
void System::get_solution(double2* h, hipStream_t st) const { P.download(h, 3ull * ndip, st); }

// ---------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------
PinnedStage::~PinnedStage() {
    if (p_) (void)hipHostFree(p_);
    for (char* q : old_) (void)hipHostFree(q);
}

// A grown stage keeps its earlier buffers until it is destroyed: freeing
// pinned memory waits for the device, which would stall a host thread that
// is preparing work while the GPU runs (the buffers add up to less than the
// last one).
char* PinnedStage::reserve(size_t bytes) {
    if (bytes > cap_) {
        if (p_) old_.push_back(p_);
        p_ = nullptr;
        const size_t c = std::max({bytes, 2 * cap_, size_t(4) << 20});
        TDDA_HIP(hipHostMalloc((void**)&p_, c, hipHostMallocDefault));
        cap_ = c;
    }
    return p_;
}

template <typename T>
void Batch::upload(DevBuf<T>& dst, const T* src, size_t n, size_t stage_off, hipStream_t st) {
    if (n == 0) return;
    if (!pinned_) {
        dst.upload(src, n, st);
        return;
    }
    char* stage = stage_.reserve(0) + stage_off;
    std::memcpy(stage, src, n * sizeof(T));
    dst.upload(reinterpret_cast<const T*>(stage), n, st);
}

// Batch buffers grow by half again their size at least: batches are rebuilt
// every few chunks with a varying number of jobs, and each new size would
// otherwise take a fresh device allocation.
template <typename T> static void grow(DevBuf<T>& b, size_t n) {
    if (b.size() < n) b.alloc(std::max<size_t>({n, b.size() + b.size() / 2, 1}));
}

void Batch::build(const std::vector<System*>& systems, KScalars* ks_dev, bool refined, hipStream_t st) {
    std::vector<KScalars*> ks(systems.size(), nullptr);
    if (ks_dev)
        for (size_t j = 0; j < systems.size(); ++j) ks[j] = ks_dev + j;
    build(systems, ks, refined, st, true);
}

int* ru_job_counter() {
    static std::mutex mu;
    static int* ptr = nullptr;
    std::lock_guard<std::mutex> lk(mu);
    if (ptr == nullptr) {
        TDDA_HIP(hipMalloc(&ptr, sizeof(int)));
        TDDA_HIP(hipMemset(ptr, 0, sizeof(int)));
    }
    return ptr;
}

namespace {

// Z tiles of one job ordered so that tiles whose pencils are x/y mirrors of
// each other (they read the same entries of the octant Green's spectrum) are
// consecutive: sort by the mirrored centre pencil of the tile.
// The order depends only on (lx, ly, tile size, tile count); batches are
// rebuilt every few chunks over the same few shapes, so each order is sorted
// once per thread and kept.
const std::vector<int>& z_tile_order(const GridJob& J, int per_tile, int ntiles) {
    static thread_local std::map<std::tuple<int, int, int, int>, std::vector<int>> cache;
    std::vector<int>& order = cache[std::make_tuple(J.lx, J.ly, per_tile, ntiles)];
    if ((int)order.size() == ntiles) return order;
    std::vector<std::pair<long, int>> key(ntiles);
    for (int t = 0; t < ntiles; ++t) {
        const long pc = (long)t * per_tile + per_tile / 2;
        const int ky = (int)(pc / J.lx), kx = (int)(pc % J.lx);
        const int my = std::min(ky, J.ly - ky), mx = std::min(kx, J.lx - kx);
        key[t] = {((long)my * J.lx + mx), t};
    }
    std::stable_sort(key.begin(), key.end(),
                     [](const std::pair<long, int>& a, const std::pair<long, int>& b) { return a.first < b.first; });
    order.resize(ntiles);
    for (int t = 0; t < ntiles; ++t) order[t] = key[t].second;
    return order;
}

// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b % 8).  Give XCD x a contiguous range of the logical order so that tiles
// adjacent in that order run together on one XCD and share its L2.
void xcd_permute(std::vector<int2>& m) {
    constexpr size_t kXcd = 8;
    const size_t n = m.size();
    if (n < 2 * kXcd) return;
    std::vector<int2> out(n);
    size_t start = 0;
    for (size_t x = 0; x < kXcd; ++x) {
        const size_t cnt = (n - x + kXcd - 1) / kXcd;
        for (size_t r = 0; r < cnt; ++r) out[x + kXcd * r] = m[start + r];
        start += cnt;
    }
    m.swap(out);
}

}  // namespace

void Batch::build(const std::vector<System*>& systems, const std::vector<KScalars*>& ks, bool refined,
                  hipStream_t st, bool interleave_shared) {
    hjobs.clear();
    std::vector<int2> m[2][NPASS][2];
    std::vector<int2> myz[NYZ];
    std::vector<int2> mxy[NXY_CLASSES];
    std::vector<int2> mu;
    const int fuse = fused_yz();
    const bool planes = plane_passes();
    size_t lds[2][NPASS][2] = {};
    int zcls[2] = {0, 0};                  // specialised Z length classes present
    size_t zlds[2][2] = {};
    for (size_t j = 0; j < systems.size(); ++j) {
        System* s = systems[j];
        TDDA_REQUIRE(!refined || s->refined, ERR_INVALID, "refined solve needs refined systems");
        GridJob J = s->job;
        J.ks = j < ks.size() ? ks[j] : nullptr;
        // c64 matvec through the fused y/z pass: x frequencies paired with
        // their mirrors (kx, lx - kx read the same octant records)
        // c64 X and Y transforms through the plane passes (one block per
        // z-plane, whose X inverse partials are per plane); they take
        // precedence over the fused y/z pass
        J.xy32 = planes && J.xyc >= 0 && schedule_plane(J.lx) ? 1 : 0;
        if (J.xy32) J.tx32 = J.nz;
        J.tr32 = !J.xy32 && J.yzc >= 0 &&
                         (fuse == FUSED_ALL || (fuse == FUSED_SCHEDULE && schedule_fused(J.ly, J.lz)))
                     ? 1
                     : 0;
        hjobs.push_back(J);
        const int jj = (int)j;
        if (J.xy32)
            for (int z = 0; z < J.nz; ++z) mxy[J.xyc].push_back(make_int2(jj, z));
        if (J.tr32) {   // 0 and lx/2 are their own mirrors
            std::vector<int2>& mz = myz[J.yzc];
            mz.push_back(make_int2(jj, 0));
            if (J.lx % 2 == 0) mz.push_back(make_int2(jj, J.lx / 2));
            for (int a = 1; 2 * a < J.lx; ++a) {
                mz.push_back(make_int2(jj, a));
                mz.push_back(make_int2(jj, J.lx - a));
            }
        }
        for (int pr = 0; pr < 2; ++pr) {
            const bool c64 = pr == 0;
            const bool fused = c64 && J.tr32;
            if (c64 && J.xy32) {   // only the Z pass runs through the pass maps
                size_t& lz = lds[pr][PASS_Z][J.fz];
                lz = std::max(lz, lds_bytes(J, PASS_Z, c64));
                if (J.fz) {
                    const int zc = z_class(J.lz, J.var[pr][PASS_Z]);
                    zcls[pr] |= 1 << zc;
                    zlds[pr][zc] = std::max(zlds[pr][zc], lds_bytes(J, PASS_Z, c64));
                }
                continue;
            }
            const int tx = c64 ? J.tx32 : J.tx64, ty = c64 ? J.ty32 : J.ty64;
            for (int t = 0; t < tx; ++t) m[pr][PASS_X][J.fx].push_back(make_int2(jj, t));
            const int txf = c64 ? J.txf32 : J.txf64;
            for (int t = 0; t < txf; ++t) m[pr][PASS_XF][J.fx].push_back(make_int2(jj, t));
            size_t& lxf = lds[pr][PASS_XF][J.fx];
            lxf = std::max(lxf, lds_bytes(J, PASS_XF, c64));
            if (fused) {
                size_t& lx = lds[pr][PASS_X][J.fx];
                lx = std::max(lx, lds_bytes(J, PASS_X, c64));
                continue;
            }
            for (int t = 0; t < ty; ++t) m[pr][PASS_Y][J.fy].push_back(make_int2(jj, t));
            size_t& lx = lds[pr][PASS_X][J.fx];
            size_t& ly = lds[pr][PASS_Y][J.fy];
            size_t& lz = lds[pr][PASS_Z][J.fz];
            lx = std::max(lx, lds_bytes(J, PASS_X, c64));
            ly = std::max(ly, lds_bytes(J, PASS_Y, c64));
            lz = std::max(lz, lds_bytes(J, PASS_Z, c64));
            if (J.fz) {
                const int zc = z_class(J.lz, J.var[pr][PASS_Z]);
                zcls[pr] |= 1 << zc;
                zlds[pr][zc] = std::max(zlds[pr][zc], lds_bytes(J, PASS_Z, c64));
            }
        }
        for (int t = 0; t < J.tupd; ++t) mu.push_back(make_int2(jj, t));
    }
    // Z maps: mirror-grouped tile order per job; systems that share an
    // operator have their tiles interleaved in chunks of 8
    std::vector<std::vector<int>> groups;
    std::map<const Operator*, size_t> gidx;
    for (size_t j = 0; j < systems.size(); ++j) {
        if (!interleave_shared) { groups.push_back({(int)j}); continue; }
        const Operator* o = systems[j]->op.get();
        auto it = gidx.find(o);
        if (it == gidx.end()) { gidx[o] = groups.size(); groups.push_back({(int)j}); }
        else groups[it->second].push_back((int)j);
    }
    for (int pr = 0; pr < 2; ++pr) {
        if (pr == 1 && refined) {
            // c128 maps of a refined batch serve only reliable updates, which
            // run over per-job compacted copies: keep each job's tiles
            // contiguous and in job order
            for (size_t j = 0; j < hjobs.size(); ++j) {
                const GridJob& J = hjobs[j];
                const std::vector<int>& order = z_tile_order(J, J.bz64, J.tz64);
                for (int t : order) m[1][PASS_Z][J.fz].push_back(make_int2((int)j, t));
            }
            continue;
        }
        for (const auto& g : groups) {
            const GridJob& J0 = hjobs[g[0]];
            if (pr == 0 && J0.tr32) continue;   // fused y/z pass (same shape for the whole group)
            const int T = pr == 0 ? J0.tz32 : J0.tz64;
            const int per = pr == 0 ? J0.bz32 : J0.bz64;
            const std::vector<int>& order = z_tile_order(J0, per, T);
            std::vector<int2>& mz = m[pr][PASS_Z][J0.fz];
            for (int t0 = 0; t0 < T; t0 += 8) {
                const int t1 = std::min(T, t0 + 8);
                for (int j : g)
                    for (int t = t0; t < t1; ++t) mz.push_back(make_int2(j, order[t]));
            }
        }
        for (int k = 0; k < 2; ++k) xcd_permute(m[pr][PASS_Z][k]);
    }
    for (auto& v : myz) xcd_permute(v);
    for (auto& v : mxy) xcd_permute(v);
    size_t total = mu.size();
    for (auto& a : m)
        for (auto& b : a)
            for (auto& v : b) total += v.size();
    for (auto& v : myz) total += v.size();
    for (auto& v : mxy) total += v.size();
    std::vector<int2>& all = hmaps_;   // kept alive: the upload is asynchronous
    all.clear();
    all.reserve(total);
    size_t off[2][NPASS][2];
    for (int pr = 0; pr < 2; ++pr)
        for (int p = 0; p < NPASS; ++p)
            for (int k = 0; k < 2; ++k) {
                off[pr][p][k] = all.size();
                all.insert(all.end(), m[pr][p][k].begin(), m[pr][p][k].end());
            }
    const size_t off_u = all.size();
    all.insert(all.end(), mu.begin(), mu.end());
    size_t off_yz[NYZ];   // after the c128 maps and the update map, which the compaction reads as one run
    for (int c = 0; c < NYZ; ++c) {
        off_yz[c] = all.size();
        all.insert(all.end(), myz[c].begin(), myz[c].end());
    }
    size_t off_xy[NXY_CLASSES];
    for (int c = 0; c < NXY_CLASSES; ++c) {
        off_xy[c] = all.size();
        all.insert(all.end(), mxy[c].begin(), mxy[c].end());
    }
    // stage layout: jobs, maps, reliable-update ranges (256-byte aligned)
    const int nj_all = (int)hjobs.size();
    auto al256 = [](size_t b) { return (b + 255) & ~size_t(255); };
    stage_jobs_ = 0;
    const size_t stage_maps = al256(hjobs.size() * sizeof(GridJob));
    const size_t stage_ranges = stage_maps + al256(total * sizeof(int2));
    if (pinned_) stage_.reserve(stage_ranges + al256((size_t)RU_REGIONS * (nj_all + 1) * sizeof(int)));
    grow(maps_, total);
    upload(maps_, all.data(), total, stage_maps, st);
    grow(djobs_, hjobs.size());
    upload_jobs(st);
    const int2* base = maps_.get();
    ls = LaunchSet{};
    ls.jobs = djobs_.get();
    ls.njobs = (int)hjobs.size();
    ls.ru_jobs = ru_job_counter();
    for (int pr = 0; pr < 2; ++pr)
        for (int p = 0; p < NPASS; ++p)
            for (int k = 0; k < 2; ++k) {
                PassMap& pm = ls.pm[pr][p][k];
                pm.map = base + off[pr][p][k];
                pm.n = (int)m[pr][p][k].size();
                pm.lds = lds[pr][p][k];
                if (p == PASS_Z && k == 1) {
                    pm.cls = zcls[pr];
                    pm.lds_cls[0] = zlds[pr][0];
                    pm.lds_cls[1] = zlds[pr][1];
                }
            }
    ls.map_upd = base + off_u;
    ls.n_upd = (int)mu.size();
    ls.refined = refined;
    for (int c = 0; c < NYZ; ++c) {
        PassMap& pm = ls.yz[c];
        pm.map = base + off_yz[c];
        pm.n = (int)myz[c].size();
        pm.lds = pm.n ? yz_lds_bytes(c) : 0;
    }
    ls.n_xy = 0;
    for (int c = 0; c < NXY_CLASSES; ++c) {
        PassMap& pm = ls.xy[c];
        pm.map = base + off_xy[c];
        pm.n = (int)mxy[c].size();
        ls.n_xy += pm.n;
    }
    // reliable-update compaction regions: the c128 maps, contiguous from
    // pm[1][X][0] in build order
    if (refined) {
        const size_t n64 = off_u - off[1][0][0];
        grow(ru_maps_, n64);
        if (ru_count_.size() < (size_t)RU_REGIONS) ru_count_.alloc(RU_REGIONS);
        ls.ru_map = ru_maps_.get();
        ls.ru_count = ru_count_.get();
        ls.ru_src = base + off[1][0][0];
        for (int p = 0; p < NPASS; ++p)
            for (int k = 0; k < 2; ++k) ls.ru_off[2 * p + k] = (int)(off[1][p][k] - off[1][0][0]);
        ls.ru_off[RU_REGIONS] = (int)n64;
        // per-region start of every job's (contiguous) tiles, [RU_REGIONS][njobs + 1]
        const int nj = (int)hjobs.size();
        std::vector<int>& rr = hranges_;   // kept alive: the upload is asynchronous
        rr.assign(RU_REGIONS * (nj + 1), 0);
        for (int p = 0; p < NPASS; ++p)
            for (int k = 0; k < 2; ++k) {
                const std::vector<int2>& v = m[1][p][k];
                int* r = rr.data() + (2 * p + k) * (nj + 1);
                size_t i = 0;
                for (int j = 0; j <= nj; ++j) {
                    while (i < v.size() && v[i].x < j) ++i;
                    r[j] = (int)i;
                }
            }
        grow(ru_ranges_, rr.size());
        upload(ru_ranges_, rr.data(), rr.size(), stage_ranges, st);
        ls.ru_ranges = ru_ranges_.get();
    }
}

void Batch::upload_jobs(hipStream_t st) { upload(djobs_, hjobs.data(), hjobs.size(), stage_jobs_, st); }

// ---------------------------------------------------------------------------
// graphs
// ---------------------------------------------------------------------------
namespace {
std::atomic<bool> g_graphs{true};

struct CaptureStream {
    hipStream_t s = nullptr;
    ~CaptureStream() {
        if (s) (void)hipStreamDestroy(s);
    }
};
}  // namespace

void set_graph_capture(bool on) { g_graphs = on; }
bool graph_capture() { return g_graphs; }

hipGraphExec_t capture_graph(const std::function<void(hipStream_t)>& enqueue) {
    static thread_local CaptureStream cs;
    if (!cs.s) TDDA_HIP(hipStreamCreateWithFlags(&cs.s, hipStreamNonBlocking));
    TDDA_HIP(hipStreamBeginCapture(cs.s, hipStreamCaptureModeThreadLocal));
    try {
        enqueue(cs.s);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(cs.s, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    hipGraph_t g = nullptr;
    TDDA_HIP(hipStreamEndCapture(cs.s, &g));
    hipGraphExec_t e = nullptr;
    const hipError_t rc = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    TDDA_HIP(rc);
    return e;
}

void launch_iteration(const LaunchSet& ls, const TwiddlePools& tw, bool refined, hipStream_t st) {
    launch_matvec(ls, tw, refined, XF_P, XI_V, GATE_ACTIVE, st);
    launch_matvec(ls, tw, refined, XF_T, XI_S, GATE_ACTIVE, st);
    if (refined) launch_matvec(ls, tw, false, XF_SOL, XI_RESID, GATE_RU, st);
}

// ---------------------------------------------------------------------------
// solving
// ---------------------------------------------------------------------------
std::vector<SolveResult> solve_batch(const std::vector<System*>& systems, const SolveOptions& opt, hipStream_t st) {
    const size_t nj = systems.size();
    std::vector<SolveResult> res(nj);
    if (nj == 0) return res;
    auto t0 = std::chrono::steady_clock::now();
    TwiddlePools tw = PlanRegistry::instance().pools();
    DevBuf<KScalars> ks(nj);
    std::vector<KScalars> hks(nj);
    std::memset(hks.data(), 0, nj * sizeof(KScalars));
    for (size_t j = 0; j < nj; ++j) {
        hks[j].tol = opt.tol;
        hks[j].maxit = opt.maxit;
        hks[j].refined = opt.refined ? 1 : 0;
        hks[j].ru_delta = opt.ru_delta;
        hks[j].tol_e = opt.refined ? opt.err_tol : 0.0;
        hks[j].status = JOB_ACTIVE;
    }
    ks.upload(hks.data(), nj, st);
    Batch b;
    b.build(systems, ks.get(), opt.refined, st);
    if (opt.warm) {
        // scratch64 = A P0 + al P0 with FP64 transforms
        for (size_t j = 0; j < nj; ++j) {
            b.hjobs[j].xin = systems[j]->P.get();
            b.hjobs[j].yout = systems[j]->scratch64.get();
        }
        b.upload_jobs(st);
        LaunchSet l64 = b.ls;
        l64.refined = opt.refined;
        launch_matvec(l64, tw, false, XF_PLAIN, XI_PLAIN, GATE_ALWAYS, st);
    }
    launch_init(b.ls, opt.warm, st);
    // preconditioned systems iterate on Y = al P
    if (opt.warm)
        for (System* s : systems)
            if (s->jacobi()) s->scale_solution(true, st);
    std::vector<KScalars> snap(nj);
    int chunk = opt.chunk > 0 ? opt.chunk : 4;
    // one iteration recorded once, replayed for every iteration
    struct ExecGuard {
        hipGraphExec_t e = nullptr;
        ~ExecGuard() {
            if (e) (void)hipGraphExecDestroy(e);
        }
    } iter_graph;
    if (graph_capture())
        iter_graph.e = capture_graph([&](hipStream_t cs) { launch_iteration(b.ls, tw, opt.refined, cs); });
    while (true) {
        for (int k = 0; k < chunk; ++k) {
            if (iter_graph.e) TDDA_HIP(hipGraphLaunch(iter_graph.e, st));
            else launch_iteration(b.ls, tw, opt.refined, st);
        }
        launch_finalize(b.ls, st);
        ks.download(snap.data(), nj, st);
        TDDA_HIP(hipStreamSynchronize(st));
        bool all_done = true;
        int max_iter = 0;
        for (size_t j = 0; j < nj; ++j) {
            if (snap[j].status == JOB_ACTIVE) all_done = false;
            max_iter = std::max(max_iter, snap[j].iter);
        }
        if (all_done) {
            for (System* s : systems)
                if (s->jacobi()) s->scale_solution(false, st);
            TDDA_HIP(hipStreamSynchronize(st));
            break;
        }
        if (opt.chunk <= 0) {
            // grow the chunk while the solve keeps going; bounded by maxit
            chunk = std::min(32, chunk * 2);
            chunk = std::max(1, std::min(chunk, opt.maxit - max_iter));
        }
    }
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    for (size_t j = 0; j < nj; ++j) {
        res[j].status = snap[j].status;
        res[j].iterations = snap[j].iterations;
        res[j].err = snap[j].err;
        res[j].n_ru = snap[j].n_ru;
        res[j].ms = ms;
        res[j].err_est = snap[j].err_est;
    }
    return res;
}

void apply_operator64(System& s, const double2* x, double2* y, hipStream_t st) {
    Batch b;
    std::vector<System*> v{&s};
    b.build(v, nullptr, false, st);
    b.hjobs[0].xin = x;
    b.hjobs[0].yout = y;
    b.upload_jobs(st);
    LaunchSet l = b.ls;
    l.refined = s.refined;
    launch_matvec(l, PlanRegistry::instance().pools(), false, XF_PLAIN, XI_PLAIN, GATE_ALWAYS, st);
    TDDA_HIP(hipStreamSynchronize(st));
}

void apply_operator32(System& s, const float2* x, float2* y, hipStream_t st) {
    TDDA_REQUIRE(s.refined, ERR_INVALID, "c64 operator needs a refined system");
    Batch b;
    std::vector<System*> v{&s};
    b.build(v, nullptr, true, st);
    b.hjobs[0].xin = x;
    b.hjobs[0].yout = y;
    b.upload_jobs(st);
    launch_matvec(b.ls, PlanRegistry::instance().pools(), true, XF_PLAIN, XI_PLAIN, GATE_ALWAYS, st);
    TDDA_HIP(hipStreamSynchronize(st));
}

void apply_batch(const std::vector<System*>& v, bool c64, const double2* x, double2* y, hipStream_t st) {
    TDDA_REQUIRE(!v.empty(), ERR_INVALID, "no systems");
    const bool refined = v[0]->refined;
    for (System* s : v) {
        TDDA_REQUIRE(s->refined == refined, ERR_INVALID, "mixed precisions in one batch");
        TDDA_REQUIRE(!c64 || s->refined, ERR_INVALID, "c64 transforms need refined systems");
    }
    Batch b;
    b.build(v, nullptr, refined, st);
    const size_t cs = c64 ? sizeof(float2) : sizeof(double2);
    std::vector<DevBuf<char>> xs(v.size()), ys(v.size());
    std::vector<std::vector<float2>> h32(v.size());
    size_t off = 0;
    for (size_t j = 0; j < v.size(); ++j) {
        const size_t n3 = 3ull * v[j]->ndip;
        xs[j].alloc(n3 * cs);
        ys[j].alloc(n3 * cs);
        if (c64) {
            h32[j].resize(n3);
            for (size_t i = 0; i < n3; ++i) h32[j][i] = make_float2((float)x[off + i].x, (float)x[off + i].y);
            TDDA_HIP(hipMemcpyAsync(xs[j].get(), h32[j].data(), n3 * cs, hipMemcpyHostToDevice, st));
        } else {
            TDDA_HIP(hipMemcpyAsync(xs[j].get(), x + off, n3 * cs, hipMemcpyHostToDevice, st));
        }
        b.hjobs[j].xin = xs[j].get();
        b.hjobs[j].yout = ys[j].get();
        off += n3;
    }
    b.upload_jobs(st);
    LaunchSet l = b.ls;
    l.refined = refined;
    launch_matvec(l, PlanRegistry::instance().pools(), c64, XF_PLAIN, XI_PLAIN, GATE_ALWAYS, st);
    off = 0;
    for (size_t j = 0; j < v.size(); ++j) {
        const size_t n3 = 3ull * v[j]->ndip;
        if (c64)
            TDDA_HIP(hipMemcpyAsync(h32[j].data(), ys[j].get(), n3 * cs, hipMemcpyDeviceToHost, st));
        else
            TDDA_HIP(hipMemcpyAsync(y + off, ys[j].get(), n3 * cs, hipMemcpyDeviceToHost, st));
        off += n3;
    }
    TDDA_HIP(hipStreamSynchronize(st));
    if (c64) {
        off = 0;
        for (size_t j = 0; j < v.size(); ++j) {
            for (size_t i = 0; i < h32[j].size(); ++i) y[off + i] = make_double2(h32[j][i].x, h32[j][i].y);
            off += h32[j].size();
        }
    }
}

void time_passes(System& s, bool c64, int reps, hipStream_t st, double* ms5) {
    std::vector<System*> v{&s};
    time_passes_batch(v, c64, reps, st, ms5);
}

void time_passes_batch(const std::vector<System*>& v, bool c64, int reps, hipStream_t st, double* ms5) {
    TDDA_REQUIRE(!v.empty(), ERR_INVALID, "no systems");
    TDDA_REQUIRE(reps >= 1, ERR_INVALID, "reps must be positive");
    const bool refined = v[0]->refined;
    for (System* s : v) {
        TDDA_REQUIRE(s->refined == refined, ERR_INVALID, "mixed precisions in one batch");
        TDDA_REQUIRE(!c64 || s->refined, ERR_INVALID, "c64 transforms need a refined system");
    }
    Batch b;
    b.build(v, nullptr, refined, st);
    std::vector<DevBuf<double2>> xs(v.size()), ys(v.size());
    for (size_t j = 0; j < v.size(); ++j) {
        const size_t n3 = 3ull * v[j]->ndip;
        xs[j].alloc(n3);
        ys[j].alloc(n3);
        xs[j].zero(st);
        b.hjobs[j].xin = xs[j].get();
        b.hjobs[j].yout = ys[j].get();
    }
    b.upload_jobs(st);
    for (int k = 0; k < 5; ++k) ms5[k] = 0.0;
    LaunchSet l = b.ls;
    launch_matvec_timed(l, PlanRegistry::instance().pools(), c64, st, 1, ms5);   // warm up
    for (int k = 0; k < 5; ++k) ms5[k] = 0.0;
    launch_matvec_timed(l, PlanRegistry::instance().pools(), c64, st, reps, ms5);
    for (int k = 0; k < 5; ++k) ms5[k] /= reps;
}

void time_iteration_batch(const std::vector<System*>& v, int reps, hipStream_t st, double* ms10) {
    TDDA_REQUIRE(!v.empty(), ERR_INVALID, "no systems");
    TDDA_REQUIRE(reps >= 1, ERR_INVALID, "reps must be positive");
    for (System* s : v) TDDA_REQUIRE(s->refined, ERR_INVALID, "iteration timing needs refined systems");
    // a real solve state: P = 0, r = r0 = rhs, then the timed steps
    DevBuf<KScalars> ks(v.size());
    std::vector<KScalars> hks(v.size());
    std::memset(hks.data(), 0, hks.size() * sizeof(KScalars));
    for (auto& k : hks) {
        k.tol = 1e-30;
        k.maxit = 1 << 30;
        k.refined = 1;
        k.ru_delta = 0.0;
    }
    ks.upload(hks.data(), hks.size(), st);
    for (System* s : v) TDDA_HIP(hipMemsetAsync(s->P.get(), 0, s->P.bytes(), st));
    Batch b;
    b.build(v, ks.get(), true, st);
    launch_init(b.ls, false, st);
    const TwiddlePools tw = PlanRegistry::instance().pools();
    double warm[5] = {0, 0, 0, 0, 0};
    launch_matvec_timed(b.ls, tw, true, st, 1, warm, XF_P, XI_V);
    launch_matvec_timed(b.ls, tw, true, st, 1, warm, XF_T, XI_S);
    for (int k = 0; k < 10; ++k) ms10[k] = 0.0;
    for (int r = 0; r < reps; ++r) {
        launch_matvec_timed(b.ls, tw, true, st, 1, ms10, XF_P, XI_V);
        launch_matvec_timed(b.ls, tw, true, st, 1, ms10 + 5, XF_T, XI_S);
    }
    for (int k = 0; k < 10; ++k) ms10[k] /= reps;
    TDDA_HIP(hipStreamSynchronize(st));
}

}  // namespace tdda
