// Host runtime: device memory, FFT plans, interaction operators, linear systems
// and the batched BiCGSTAB driver.
#pragma once

#include <condition_variable>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <vector>

#include "dda_types.h"
#include "green.h"
#include "passes.h"

namespace tdda {

// ---------------------------------------------------------------------------
// device buffer (move-only RAII)
// ---------------------------------------------------------------------------
size_t device_bytes_in_use();     // live blocks (rounded sizes)
size_t device_bytes_cached();     // freed blocks kept for reuse
void device_release_cached();     // hipFree every cached block

template <typename T> class DevBuf {
   public:
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
        return *this;
    }
    void alloc(size_t n);
    void release();
    T* get() const { return p_; }
    size_t size() const { return n_; }
    size_t bytes() const { return n_ * sizeof(T); }
    void upload(const T* h, size_t n, hipStream_t st) {
        TDDA_HIP(hipMemcpyAsync(p_, h, n * sizeof(T), hipMemcpyHostToDevice, st));
    }
    void download(T* h, size_t n, hipStream_t st) const {
        TDDA_HIP(hipMemcpyAsync(h, p_, n * sizeof(T), hipMemcpyDeviceToHost, st));
    }
    void zero(hipStream_t st) { if (n_) TDDA_HIP(hipMemsetAsync(p_, 0, n_ * sizeof(T), st)); }

   private:
    T* p_ = nullptr;
    size_t n_ = 0;
};

void* dev_alloc(size_t bytes);
void dev_free(void* p, size_t bytes);

template <typename T> void DevBuf<T>::alloc(size_t n) {
    release();
    if (n == 0) return;
    p_ = static_cast<T*>(dev_alloc(n * sizeof(T)));
    n_ = n;
}
template <typename T> void DevBuf<T>::release() {
    if (p_) dev_free(p_, n_ * sizeof(T));
    p_ = nullptr;
    n_ = 0;
}

// ---------------------------------------------------------------------------
// FFT lengths, plans and twiddle tables
// ---------------------------------------------------------------------------
constexpr int MAX_LXY = 256;   // limits set by the per-tile caps (3*Lx <= X cap in FP64)
constexpr int MAX_LZ = 256;

int choose_fft_length(int n);             // smallest 2,3,5-smooth L >= max(2, 2n-1)
std::vector<int> radix_sequence(int L);   // forward stage radices from {8,4,2,5,3}

class PlanRegistry {
   public:
    static PlanRegistry& instance();
    FftPlan plan(int L);   // builds and uploads twiddles on first use (synchronous)
    TwiddlePools pools();

   private:
    std::mutex mu_;
    std::map<int, FftPlan> plans_;
    DevBuf<float2> tw32_;
    DevBuf<double2> tw64_;
    int used_ = 0;
};

// ---------------------------------------------------------------------------
// Si/Ci tables (device copies keyed by content)
// ---------------------------------------------------------------------------
struct SiciDevice {
    const double* si = nullptr;
    const double* ci = nullptr;
    long n = 0;
    double delta = 0.1;
    uint64_t id = 0;
};
SiciDevice sici_tables(const double* si, const double* ci, long n, double delta);

// ---------------------------------------------------------------------------
// interaction operator: octant Green's spectrum for one (grid, K, images, method)
// ---------------------------------------------------------------------------
struct OperatorKey {
    int nx, ny, nz;
    double d, K;
    int method, maxm, maxn;
    double lm, ln;
    uint64_t sici_id;
    bool operator<(const OperatorKey& o) const;
};

// Grow-only pinned host buffer: asynchronous uploads from it do not wait for
// the stream (uploads from pageable memory block the host until the stream
// reaches them).  The owner must not overwrite bytes an upload still reads.
class PinnedStage {
   public:
    PinnedStage() = default;
    PinnedStage(const PinnedStage&) = delete;
    PinnedStage& operator=(const PinnedStage&) = delete;
    ~PinnedStage();
    char* reserve(size_t bytes);   // contents are not kept on growth
   private:
    char* p_ = nullptr;
    size_t cap_ = 0;
    std::vector<char*> old_;
};

class GreenBatch;

// The build is enqueued on the caller's stream and not waited for: work on
// that stream sees the spectrum; other streams wait on ready() (the cache
// does this on every hit).  The build's temporaries live until it is done.
// With a GreenBatch the steps are queued in the batch instead (a lattice sum
// with images is still launched at once) and launched by its flush().
class Operator {
   public:
    Operator(const OperatorKey& key, const SiciDevice& tables, hipStream_t st, GreenBatch* batch = nullptr);
    ~Operator();
    Operator(const Operator&) = delete;
    Operator& operator=(const Operator&) = delete;
    OperatorKey key;
    int nx, ny, nz, lx, ly, lz, hx, hy, hz;
    FftPlan px, py, pz;
    DevBuf<double2> green64;   // [hz][hy][6][hx] (c64: [hz][hy][3][hx] of component pairs)
    DevBuf<float2> green32;
    size_t bytes() const { return green64.bytes() + green32.bytes(); }
    hipEvent_t ready() const { return done_; }
    // GPU time of the build (lattice sum, DFTs, layout) in ms; waits for it
    // (-1 for an operator built in a batch: its steps ran with the others')
    double setup_ms() const;
    // free the build temporaries once the build has run; true when none are left
    bool reap();

   private:
    friend class GreenBatch;
    GreenParams gp_{};
    const double *si_ = nullptr, *ci_ = nullptr;
    hipEvent_t start_ = nullptr, done_ = nullptr;
    std::vector<DevBuf<double2>> tmp_;
};

// Green's set-ups of a round of new operators in one launch per step (octant
// tensor, three DFT axes, layout) instead of five launches per operator.
// Operators queued in the batch are published to the cache by flush(); until
// then their keys count as being built, and the cache hands later requests
// for the same key in the same batch the queued operator.  A batch is reused
// round after round (a flush first waits for the previous flush's launches,
// which read the job tables it overwrites).
class GreenBatch {
   public:
    GreenBatch() = default;
    GreenBatch(const GreenBatch&) = delete;
    GreenBatch& operator=(const GreenBatch&) = delete;
    ~GreenBatch();
    void add(const std::shared_ptr<Operator>& op, hipStream_t st);
    std::shared_ptr<Operator> find(const OperatorKey& key) const;
    void flush(hipStream_t st);
    // drop the queued operators without building them (after a failure):
    // their keys are released and they never enter the cache
    void cancel();
    bool empty() const { return ops_.empty(); }

   private:
    std::vector<std::shared_ptr<Operator>> ops_;
    hipStream_t st_ = nullptr;   // stream of the queued lattice sums
    PinnedStage stage_;
    DevBuf<char> dev_;
    hipEvent_t flushed_ = nullptr;   // recorded after the last flush's launches
};

// LRU cache of operators bounded by bytes
class OperatorCache {
   public:
    static OperatorCache& instance();
    // batch: queue a missing operator's set-up there (published by its flush)
    std::shared_ptr<Operator> get(const OperatorKey& key, const SiciDevice& tables, hipStream_t st, bool* hit,
                                  GreenBatch* batch = nullptr);
    void set_budget(size_t bytes) { budget_ = bytes; }
    size_t bytes() const;
    void clear();
    // evict least recently used operators nobody holds until `bytes` are freed;
    // returns the bytes freed
    size_t trim(size_t bytes);

   private:
    friend class GreenBatch;
    // a batch's operators enter the cache (built: their ready() events are
    // recorded) or, abandoned, give up their keys
    void publish(const std::vector<std::shared_ptr<Operator>>& ops, bool built);
    void insert_locked(const OperatorKey& key, const std::shared_ptr<Operator>& op);
    void touch_locked(const OperatorKey& key);   // move key to the most recent end of lru_
    void reap_locked();                          // free finished builds' temporaries
    mutable std::mutex mu_;
    std::condition_variable built_;
    std::map<OperatorKey, std::shared_ptr<Operator>> map_;
    std::set<OperatorKey> building_;   // keys one thread is building (others wait for it)
    std::vector<std::weak_ptr<Operator>> unreaped_;   // builds whose temporaries are still held
    std::vector<OperatorKey> lru_;     // each cached key once, least recent first
    size_t budget_ = size_t(8) << 30;
};

// ---------------------------------------------------------------------------
// one linear system (diag(al) + A) P = rhs and its device memory
// ---------------------------------------------------------------------------
class System {
   public:
    // box_map: [nz][ny][nx] -> dipole index or -1; null means the full box in
    // x-fastest order (python_bindings/dda_model/__init__.py:8-20).
    System(std::shared_ptr<Operator> op, int ndip, const int* box_map_host, bool refined, hipStream_t st);
    void set_alpha(const double2* al_dip_host, hipStream_t st);   // 1/alpha per dipole
    void set_alpha_device(const double2* al_dip_dev, hipStream_t st);
    // local-field factor f = al 3/(eps + 2) per dipole (refined systems): the
    // field-error estimate weighs the solution by it (E = f P)
    void set_field_factor(const double2* f_dip_host, hipStream_t st);
    // right Jacobi preconditioning of later solves (off: the reference's plain
    // BiCGSTAB, DDAModel.cpp:899-1009)
    void set_jacobi(bool on) { job.jacobi = on ? 1 : 0; }
    bool jacobi() const { return job.jacobi != 0; }
    // P <- P / al (Y of a preconditioned solve back to the polarisation) or
    // P <- al P (a polarisation to Y)
    void scale_solution(bool to_y, hipStream_t st);
    void set_rhs(const double2* rhs_host, hipStream_t st);
    void set_solution(const double2* P_host, hipStream_t st);
    void get_solution(double2* P_host, hipStream_t st) const;
    size_t bytes() const;
    static size_t estimate_bytes(const Operator& op, int ndip, bool refined);

    std::shared_ptr<Operator> op;
    int ndip;
    bool refined;
    GridJob job;   // host copy of the descriptor (ks set by the batch)
    DevBuf<char> grid;          // c128-sized work grid (c64 view aliases it)
    DevBuf<char> kvec;          // r, r0, p, v, t, s (+ dP in refined mode)
    DevBuf<double2> P, rhs, scratch64;
    DevBuf<double2> al64, ial64;   // 1/alpha and alpha per dipole
    DevBuf<float2> al32, ial32;
    DevBuf<double> partial;     // X inverse partials, then (refined) the X forward ones
    DevBuf<float2> efac;        // refined: local-field factor per dipole
    DevBuf<int> box_map;
};

struct SolveOptions {
    bool refined = false;   // c64 Krylov + FP64 reliable updates (requires refined systems)
    int maxit = 100000;
    double tol = 1e-5;
    double ru_delta = 0.01;   // reliable update when the residual falls this far below its peak since the last one
    // refined mode: relative field-error target (estimated from the solve
    // itself, see KScalars); 0: stop on the true residual alone
    double err_tol = 0.0;
    bool warm = false;      // use the system's current P as the initial guess
    int chunk = 0;          // iterations between host checks (0: adaptive)
};

// Graph capture of launch sequences (on by default): the sequence is
// recorded once on a thread-local side stream and replayed with one graph
// launch.  Off: the same launches are issued one by one.
void set_graph_capture(bool on);
bool graph_capture();
// Records enqueue(stream) into an executable graph (thread-local capture mode).
hipGraphExec_t capture_graph(const std::function<void(hipStream_t)>& enqueue);
// One BiCGSTAB step of a batch as a graph: the p and t matvecs and, in refined
// mode, the reliable-update matvec (gated per job on the device).
void launch_iteration(const LaunchSet& ls, const TwiddlePools& tw, bool refined, hipStream_t st);

struct SolveResult {
    int status = JOB_IDLE;
    int iterations = 0;
    double err = 0.0;
    int n_ru = 0;
    double ms = 0.0;
    double err_est = 0.0;   // refined mode: estimated relative field error (0: none)
};


// Device-resident batch: descriptors and (job, tile) maps for every pass.
class Batch {
   public:
    // Upload through a pinned stage (batches rebuilt while the GPU runs:
    // the stage of a build must not be rebuilt before its uploads ran, which
    // the server guarantees by alternating two batches behind events).
    void set_pinned(bool on) { pinned_ = on; }
    // ks_dev: scalars of system j at ks_dev + j (or null for plain operator use)
    void build(const std::vector<System*>& systems, KScalars* ks_dev, bool refined, hipStream_t st);
    // explicit scalar slot per system; systems that share an Operator get
    // their Z-pass tiles interleaved so that co-resident blocks read the same
    // Green's spectrum (blocks congruent mod 8 run on one XCD and share its L2)
    void build(const std::vector<System*>& systems, const std::vector<KScalars*>& ks, bool refined,
               hipStream_t st, bool interleave_shared = true);
    void upload_jobs(hipStream_t st);   // after editing hjobs (xin/yout)
    LaunchSet ls;
    std::vector<GridJob> hjobs;

   private:
    template <typename T> void upload(DevBuf<T>& dst, const T* src, size_t n, size_t stage_off, hipStream_t st);
    DevBuf<GridJob> djobs_;
    DevBuf<int2> maps_, ru_maps_;
    DevBuf<int> ru_count_, ru_ranges_;
    std::vector<int> hranges_;
    std::vector<int2> hmaps_;
    bool pinned_ = false;
    PinnedStage stage_;
    size_t stage_jobs_ = 0;   // offsets of the regions in the stage
};

// Solve every system in the batch with BiCGSTAB (DDAModel.cpp:899-1009).
std::vector<SolveResult> solve_batch(const std::vector<System*>& systems, const SolveOptions& opt, hipStream_t st);

// y = A x + al x with FP64 transforms (x, y device, 3*ndip)
void apply_operator64(System& s, const double2* x, double2* y, hipStream_t st);
// same with c64 transforms (requires a refined system)
void apply_operator32(System& s, const float2* x, float2* y, hipStream_t st);
// y_j = A_j x_j + al_j x_j for every system in one batched launch per pass;
// x, y: host vectors of the systems back to back (3 ndip_j complex each);
// c64 transforms need refined systems
void apply_batch(const std::vector<System*>& v, bool c64, const double2* x, double2* y, hipStream_t st);

// Average ms per launch of the five passes (X fwd, Y fwd, Z, Y inv, X inv)
// of y = A x + al x over `reps` repetitions, c64 or c128 transforms (plane
// jobs: x+y forward in slot 0, inverse y+x in slot 4; empty slots read 0).
void time_passes(System& s, bool c64, int reps, hipStream_t st, double* ms5);
// same over a batch of systems in one launch per pass
void time_passes_batch(const std::vector<System*>& v, bool c64, int reps, hipStream_t st, double* ms5);
// Average ms per launch of the ten passes of one refined BiCGSTAB step
// (X fwd with p update ... X inv with <r0,v>, then X fwd with t ... X inv with
// the XI_S dots) over a batch, on zeroed Krylov state.
void time_iteration_batch(const std::vector<System*>& v, int reps, hipStream_t st, double* ms10);

hipStream_t default_stream();

}  // namespace tdda
