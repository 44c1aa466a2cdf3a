// Request server: continuous batching of DDA solves of uneven shape on one GPU.
//
// Requests (box shape, wavelength, incidence, density) are queued by submit().
// An admission thread takes them in arrival order while a slot is free and
// the device-memory estimate fits the budget, and builds (or reuses from the
// cache) their Green's spectra and per-dipole data on a setup stream.  A loop
// thread runs the compute stream: requests join its batch once their set-up
// has completed, and every joined solve advances together (each launch of
// the five FFT passes and the Krylov kernels covers all active requests).
// Finished requests leave the batch after every chunk of iterations; their
// slots and device buffers are reused by the next arrivals.
#pragma once

#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>

#include "problem.h"

namespace tdda {

struct ServerOptions {
    bool refined = true;             // c64 Krylov + FP64 reliable updates; false: FP64
    int maxit = 100000;
    double tol = 1e-5;
    double ru_delta = 0.01;   // reliable update when the residual falls this far below its peak since the last one
    double err_tol = 1e-5;           // refined: field-error target (0: true residual only)
    bool jacobi = false;             // right Jacobi preconditioning of every solve
    size_t hbm_budget = size_t(200) << 30;
    int max_slots = 64;
    int chunk = 4;                   // iterations between host checks
    int admit_per_round = 32;        // largest admission round (rounds start at 8 and grow with the live count)
    bool keep_fields = true;         // keep E for take_field()
};

struct ServeRequest {
    long id = 0;
    ProblemDesc desc{};
    std::vector<double> rho;
    double t_submit = 0.0;
};

struct ServeResult {
    long id = 0;
    int status = JOB_IDLE;
    int iterations = 0;
    int n_ru = 0;
    double err = 0.0;
    double err_est = 0.0;
    double t_submit = 0.0, t_admit = 0.0, t_done = 0.0;   // seconds since server start
    int op_hit = 0;
};

struct ServerStats {
    long submitted = 0, completed = 0, chunks = 0, rebuilds = 0, launches = 0;
    int max_active = 0;
    size_t peak_bytes = 0;
    // host ms: admission thread (admit), loop thread (wait for a chunk,
    // retire, join and rebuild)
    double host_ms[4] = {0, 0, 0, 0};
    double enqueue_ms = 0;     // loop thread: enqueueing chunks (launches or graph replays)
    long graph_chunks = 0;     // chunks replayed from a captured graph
};

class Server {
   public:
    Server(const ServerOptions& o, const SiciDevice& tables);
    ~Server();
    void submit(ServeRequest&& r);
    int poll(std::vector<ServeResult>& out, int max);
    // 0: moved out; 1: no field kept for id; 2: the field has another length (kept)
    int take_field(long id, size_t n, std::vector<double2>& E);
    bool idle();
    double now() const;
    ServerStats stats();
    std::string error();   // first loop failure, if any

   private:
    struct Slot {
        bool used = false;
        unsigned gen = 0;   // bumped at every admission into the slot
        std::unique_ptr<Problem> prob;
        ServeRequest req;
        double t_admit = 0.0;
        int op_hit = 0;
    };
    void loop();
    void admit_loop();
    std::vector<int> admit();
    void admit_requests(int limit, GreenBatch& gb, std::vector<int>& joined);
    void fail(const std::string& what);
    void retire(const std::vector<int>& slots, const KScalars* snap);
    // finished requests whose field is still being copied to the host
    struct Retiring {
        ServeResult r;
        std::unique_ptr<Problem> prob;   // kept until the copy has read its buffers
        std::tuple<int, int, int> shape;
        DevBuf<double2> dev;
        double2* host = nullptr;         // pinned
        size_t n = 0;
        hipEvent_t ev = nullptr;
    };
    std::deque<Retiring> retiring_;
    std::map<size_t, std::vector<double2*>> pinned_fields_;   // pinned buffers by length
    void complete_retirements(bool wait);
    void finish(Retiring& rt, double t);

    ServerOptions opt_;
    SiciDevice tables_;
    std::chrono::steady_clock::time_point t0_;
    std::thread th_, adm_th_;
    // mu_ guards the queues, results, statistics, the slots' used flags and
    // generations, the join rounds and the event pool; cv_ wakes the loop
    // thread (a join round is ready or stop), cv_adm_ the admission thread
    // (a request arrived, a slot or memory was freed, or stop)
    std::mutex mu_;
    std::condition_variable cv_, cv_adm_;
    std::deque<ServeRequest> pending_;
    std::deque<ServeResult> done_;
    std::map<long, std::vector<double2>> fields_;
    bool stop_ = false;
    std::string error_;
    ServerStats stats_;

    // loop-thread state.  Two batches alternate: while the chunk of one runs
    // on the GPU, the host retires what the previous chunk finished, joins
    // admitted requests and builds the other batch for the next chunk.
    hipStream_t main_ = nullptr, setup_ = nullptr;
    DevBuf<KScalars> ks_;
    KScalars* snap_[2] = {nullptr, nullptr};   // pinned copies of ks_ after each chunk
    hipEvent_t chunk_done_[2] = {nullptr, nullptr};   // chunk of batch i complete (snap_[i] valid)
    bool inflight_[2] = {false, false};
    std::vector<Slot> slots_;
    std::vector<std::pair<int, unsigned>> members_[2];   // (slot, generation) in batch i
    // A batch holds every request admitted when it was built (finished ones
    // stay in it, gated off, until it is rebuilt).  Admissions bump
    // live_version_; a batch is rebuilt before use when its version is older or
    // more than half its members have finished.  A batch used a fourth time
    // without a rebuild has its chunk recorded as a graph and replayed.
    long live_version_ = 0;
    long batch_version_[2] = {-1, -1};
    int batch_uses_[2] = {0, 0};
    hipGraphExec_t chunk_graph_[2] = {nullptr, nullptr};
    bool enqueue_chunk(int b);   // true: replayed from the batch's graph
    void drop_graph(int b);
    Batch batch_[2], init_batch_[2];
    // Admitted requests whose operators and densities are still being set up
    // on the setup stream, by admission round.  A round joins the batch once
    // its event has completed, so the compute stream never waits for set-up
    // while other requests can run.
    struct JoinRound {
        std::vector<int> slots;
        hipEvent_t ev = nullptr;
    };
    std::deque<JoinRound> joining_;
    std::vector<hipEvent_t> free_events_;
    // slots of the rounds whose set-up is done (waits for the oldest when
    // nothing else could run)
    std::vector<int> ready_joiners(bool wait);
    // admission-thread state: densities of an admission round go up through
    // a pinned stage; two stages alternate, each reused after the round two
    // back has uploaded
    RoundSetup round_setup_[2];   // density upload and set-up of a round (alternating)
    GreenBatch green_batch_[2];   // Green's set-ups of a round's new operators (alternating)
    hipEvent_t rho_done_[2] = {nullptr, nullptr};
    bool rho_pending_[2] = {false, false};
    int rho_round_ = 0;
    // systems of finished requests by box shape, reused by admissions
    std::mutex pool_mu_;
    std::map<std::tuple<int, int, int>, std::vector<std::unique_ptr<System>>> pool_;
    int busy_ = 0;   // requests admitted but not finished (guarded by mu_)
};

}  // namespace tdda
