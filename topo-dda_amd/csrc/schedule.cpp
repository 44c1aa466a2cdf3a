// Kernel schedule and autotuner (see schedule.h).
#include "schedule.h"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <random>
#include <sstream>

#include "fft_fast.h"
#include "passes.h"

namespace tdda {

namespace {

// A tuned choice leaves the default (variant 0, no fused y/z pass, plane
// passes on) only when the alternative is faster by this factor: timings of
// a few repetitions differ by a few per cent from run to run, and near-ties
// would otherwise flip between runs.
constexpr double kTuneMargin = 0.97;
// A Z-pass variant in another kernel class than variant 0 (z_class) needs a
// wider margin: in a batch of mixed boxes the Z pass then runs as two
// launches, and the per-shape timing does not carry over (8 x 6 at Lz = 48
// won the per-shape timing but served 4.7 % fewer solves/s than 12 x 4 in a
// same-box A/B of the mixed stream).
constexpr double kClassMargin = 0.85;

const char* const kPrec[2] = {"c64", "c128"};
const char* const kPass[NPASS] = {"xinv", "y", "z", "xfwd"};

int name_index(const char* const* names, int n, const std::string& s) {
    for (int i = 0; i < n; ++i)
        if (s == names[i]) return i;
    return -1;
}

void check_variant(int pr, int pass, int L, int v) {
    TDDA_REQUIRE(pr == 0 || pr == 1, ERR_INVALID, "schedule precision must be c64 or c128");
    TDDA_REQUIRE(pass >= 0 && pass < NPASS, ERR_INVALID, "schedule pass must be xinv, y, z or xfwd");
    TDDA_REQUIRE(L >= 1, ERR_INVALID, "schedule length must be positive");
    TDDA_REQUIRE(v == 0 || (v == 1 && split_variants(pass, L) > 1), ERR_INVALID,
                 "variant " + std::to_string(v) + " does not exist for " + kPass[pass] + " at length " +
                     std::to_string(L));
}

}  // namespace

Schedule& Schedule::instance() {
    static Schedule s;
    return s;
}

Schedule::Schedule() {
    const char* path = std::getenv("TDDA_SCHEDULE");
    if (path == nullptr || *path == 0) return;
    try {
        load(path);
    } catch (const std::exception& e) {
        // the defaults are correct, only possibly slower
        std::fprintf(stderr, "topodda: schedule %s not used: %s\n", path, e.what());
    }
}

int Schedule::variant(int pr, int pass, int L) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = var_.find(std::make_tuple(pr, pass, L));
    return it == var_.end() ? 0 : it->second;
}

bool Schedule::fused(int ly, int lz) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = fused_.find(std::make_pair(ly, lz));
    return it != fused_.end() && it->second != 0;
}

bool Schedule::plane(int L) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = plane_.find(L);
    return it == plane_.end() || it->second != 0;
}

void Schedule::set_plane(int L, bool on) {
    TDDA_REQUIRE(L >= 1, ERR_INVALID, "schedule length must be positive");
    std::lock_guard<std::mutex> lk(mu_);
    plane_[L] = on ? 1 : 0;
}

void Schedule::set_variant(int pr, int pass, int L, int v) {
    check_variant(pr, pass, L, v);
    std::lock_guard<std::mutex> lk(mu_);
    var_[std::make_tuple(pr, pass, L)] = v;
}

void Schedule::set_fused(int ly, int lz, bool on) {
    TDDA_REQUIRE(ly >= 1 && lz >= 1, ERR_INVALID, "schedule lengths must be positive");
    std::lock_guard<std::mutex> lk(mu_);
    fused_[std::make_pair(ly, lz)] = on ? 1 : 0;
}

void Schedule::clear() {
    std::lock_guard<std::mutex> lk(mu_);
    var_.clear();
    fused_.clear();
    plane_.clear();
}

std::string Schedule::text() {
    std::lock_guard<std::mutex> lk(mu_);
    std::ostringstream o;
    o << "# topodda schedule\n";
    for (const auto& kv : var_)
        o << "var " << kPrec[std::get<0>(kv.first)] << " " << kPass[std::get<1>(kv.first)] << " "
          << std::get<2>(kv.first) << " " << kv.second << "\n";
    for (const auto& kv : fused_) o << "yz " << kv.first.first << " " << kv.first.second << " " << kv.second << "\n";
    for (const auto& kv : plane_) o << "xy " << kv.first << " " << kv.second << "\n";
    return o.str();
}

void Schedule::parse(const std::string& text) {
    std::map<std::tuple<int, int, int>, int> var;
    std::map<std::pair<int, int>, int> fused;
    std::map<int, int> plane;
    std::istringstream in(text);
    std::string line;
    int lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        std::istringstream ls(line);
        std::string kind;
        if (!(ls >> kind)) continue;   // blank or comment
        const std::string where = "schedule line " + std::to_string(lineno) + ": ";
        std::string extra;
        if (kind == "var") {
            std::string prec, pass;
            int L = 0, v = -1;
            TDDA_REQUIRE((ls >> prec >> pass >> L >> v) && !(ls >> extra), ERR_INVALID,
                         where + "expected 'var <c64|c128> <xinv|y|z|xfwd> <L> <variant>'");
            const int pr = name_index(kPrec, 2, prec), ps = name_index(kPass, NPASS, pass);
            TDDA_REQUIRE(pr >= 0 && ps >= 0, ERR_INVALID, where + "unknown precision or pass");
            try {
                check_variant(pr, ps, L, v);
            } catch (const Error& e) {
                throw Error(ERR_INVALID, where + e.what());
            }
            var[std::make_tuple(pr, ps, L)] = v;
        } else if (kind == "yz") {
            int ly = 0, lz = 0, on = -1;
            TDDA_REQUIRE((ls >> ly >> lz >> on) && !(ls >> extra) && ly >= 1 && lz >= 1 && (on == 0 || on == 1),
                         ERR_INVALID, where + "expected 'yz <Ly> <Lz> <0|1>'");
            fused[std::make_pair(ly, lz)] = on;
        } else if (kind == "xy") {
            int L = 0, on = -1;
            TDDA_REQUIRE((ls >> L >> on) && !(ls >> extra) && L >= 1 && (on == 0 || on == 1), ERR_INVALID,
                         where + "expected 'xy <L> <0|1>'");
            plane[L] = on;
        } else {
            throw Error(ERR_INVALID, where + "unknown entry '" + kind + "'");
        }
    }
    std::lock_guard<std::mutex> lk(mu_);
    var_.swap(var);
    fused_.swap(fused);
    plane_.swap(plane);
}

void Schedule::load(const std::string& path) {
    std::ifstream f(path);
    TDDA_REQUIRE(f.good(), ERR_INVALID, "cannot read schedule file " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    parse(ss.str());
}

void Schedule::save(const std::string& path) {
    const std::string t = text();
    std::ofstream f(path);
    TDDA_REQUIRE(f.good(), ERR_INVALID, "cannot write schedule file " + path);
    f << t;
    TDDA_REQUIRE(f.good(), ERR_INVALID, "writing schedule file " + path + " failed");
}

int schedule_variant(int pr, int pass, int L) { return Schedule::instance().variant(pr, pass, L); }
bool schedule_fused(int ly, int lz) { return Schedule::instance().fused(ly, lz); }
bool schedule_plane(int L) { return Schedule::instance().plane(L); }

// ---------------------------------------------------------------------------
// tuner
// ---------------------------------------------------------------------------
TuneReport tune(const std::vector<std::array<int, 3>>& boxes, int count, int reps, hipStream_t st) {
    TDDA_REQUIRE(!boxes.empty() && count >= 1 && reps >= 1, ERR_INVALID, "tune needs boxes, count >= 1 and reps >= 1");
    for (const auto& b : boxes)
        TDDA_REQUIRE(b[0] >= 1 && b[1] >= 1 && b[2] >= 1, ERR_INVALID, "tune: empty box");
    Schedule& S = Schedule::instance();
    struct RestoreMode {
        int mode;
        bool planes;
        ~RestoreMode() {
            set_fused_yz(mode);
            set_plane_passes(planes);
        }
    } restore{fused_yz(), plane_passes()};
    std::mt19937_64 rng(20241204);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    const SiciDevice none{};   // LDR: no tables
    auto make = [&](const std::array<int, 3>& b, bool refined) {
        std::vector<std::unique_ptr<Problem>> ps;
        for (int i = 0; i < count; ++i) {
            ProblemDesc d{};
            d.nx = b[0];
            d.ny = b[1];
            d.nz = b[2];
            d.d = 15.0;
            d.lam = 400.0 + 400.0 * U(rng);
            d.nback = 1.0;
            d.method = 1;
            d.nk[2] = 1.0;
            d.ne0[0] = 1.0;
            d.e0 = 1.0;
            d.mat[0] = 1.0;
            d.mat[2] = 5.96;
            d.mat[3] = 3.89;
            auto p = std::make_unique<Problem>(d, none, refined, st);
            std::vector<double> rho(p->ndip);
            for (double& x : rho) x = U(rng);
            p->set_density(rho.data(), false, st);
            TDDA_HIP(hipStreamSynchronize(st));
            ps.push_back(std::move(p));
        }
        return ps;
    };
    auto time5 = [&](const std::array<int, 3>& b, bool c64, double* ms5) {
        auto ps = make(b, c64);
        std::vector<System*> sys;
        for (auto& p : ps) sys.push_back(p->sys.get());
        time_passes_batch(sys, c64, reps, st, ms5);
    };
    auto lengths = [](const std::array<int, 3>& b, int* L) {   // per PassKind
        const int lx = choose_fft_length(b[0]), ly = choose_fft_length(b[1]), lz = choose_fft_length(b[2]);
        L[PASS_X] = lx;
        L[PASS_Y] = ly;
        L[PASS_Z] = lz;
        L[PASS_XF] = lx;
    };

    TuneReport rep;
    // 1. split variants of every pass, with the fused and plane passes off
    //    (c64 X, Y and Z timings are those of the separate passes)
    set_fused_yz(FUSED_OFF);
    set_plane_passes(false);
    for (int pr = 0; pr < 2; ++pr) {
        std::map<std::pair<int, int>, std::array<double, 2>> acc;   // (pass, L) -> ms per variant
        for (int v = 0; v < 2; ++v) {
            bool any = false;
            for (const auto& b : boxes) {
                int L[NPASS];
                lengths(b, L);
                for (int p = 0; p < NPASS; ++p)
                    if (split_variants(p, L[p]) > 1) {
                        S.set_variant(pr, p, L[p], v);
                        any = true;
                    }
            }
            if (v == 1 && !any) break;
            for (const auto& b : boxes) {
                int L[NPASS];
                lengths(b, L);
                double ms5[5];
                time5(b, pr == 0, ms5);
                // launch order: X forward, Y forward, Z, Y inverse, X inverse
                const double per_pass[NPASS] = {ms5[4], ms5[1] + ms5[3], ms5[2], ms5[0]};
                for (int p = 0; p < NPASS; ++p) {
                    if (v == 1 && split_variants(p, L[p]) < 2) continue;
                    auto it = acc.find({p, L[p]});
                    if (it == acc.end()) it = acc.emplace(std::make_pair(p, L[p]), std::array<double, 2>{0.0, -1.0}).first;
                    if (v == 1 && it->second[1] < 0) it->second[1] = 0.0;
                    it->second[v] += per_pass[p];
                }
            }
        }
        for (const auto& kv : acc) {
            const int p = kv.first.first, L = kv.first.second;
            const double margin =
                (p == PASS_Z && z_class(L, 0) != z_class(L, 1)) ? kClassMargin : kTuneMargin;
            const int chosen = (kv.second[1] >= 0 && kv.second[1] < margin * kv.second[0]) ? 1 : 0;
            if (split_variants(p, L) > 1) S.set_variant(pr, p, L, chosen);
            rep.passes.push_back(TunePass{pr, p, L, {kv.second[0], kv.second[1]}, chosen});
        }
    }
    // 2. the fused y/z pass against the three passes (whole c64 matvec, the
    //    chosen variants)
    std::map<std::pair<int, int>, std::array<double, 2>> facc;
    for (const auto& b : boxes) {
        int L[NPASS];
        lengths(b, L);
        int per = 0;
        size_t bytes = 0;
        const bool fast_x = fast_tiles(L[PASS_X], PASS_X, true, 0, &per, &bytes);
        if (!fast_x || yz_class(L[PASS_Y], L[PASS_Z], b[1], b[2]) < 0) continue;
        auto& a = facc[{L[PASS_Y], L[PASS_Z]}];
        for (int f = 0; f < 2; ++f) {
            set_fused_yz(f ? FUSED_ALL : FUSED_OFF);
            double ms5[5];
            time5(b, true, ms5);
            a[f] += ms5[0] + ms5[1] + ms5[2] + ms5[3] + ms5[4];
        }
    }
    for (const auto& kv : facc) {
        const bool on = kv.second[1] < kTuneMargin * kv.second[0];
        S.set_fused(kv.first.first, kv.first.second, on);
        rep.fused.push_back(TuneFused{kv.first.first, kv.first.second, kv.second[0], kv.second[1], on});
    }
    // 3. the plane passes per plane length against the choices above (the
    //    y/z choice as recorded, plane passes off), whole c64 matvec
    set_fused_yz(FUSED_SCHEDULE);
    std::map<int, std::array<double, 2>> pacc;
    for (const auto& b : boxes) {
        int L[NPASS];
        lengths(b, L);
        if (plane_class_of(L[PASS_X], L[PASS_Y], b[0], b[1]) < 0) continue;
        auto& a = pacc[L[PASS_X]];
        for (int on = 0; on < 2; ++on) {
            set_plane_passes(on != 0);
            S.set_plane(L[PASS_X], on != 0);
            double ms5[5];
            time5(b, true, ms5);
            a[on] += ms5[0] + ms5[1] + ms5[2] + ms5[3] + ms5[4];
        }
    }
    for (const auto& kv : pacc) {
        const bool on = !(kv.second[0] < kTuneMargin * kv.second[1]);
        S.set_plane(kv.first, on);
        rep.planes.push_back(TunePlane{kv.first, kv.second[0], kv.second[1], on});
    }
    return rep;
}

std::string tune_report_text(const TuneReport& r) {
    std::ostringstream o;
    char buf[160];
    for (const auto& p : r.passes) {
        if (p.ms[1] >= 0)
            std::snprintf(buf, sizeof buf, "pass %s %s %d ms %.7f %.7f chosen %d\n", kPrec[p.pr], kPass[p.pass], p.L,
                          p.ms[0], p.ms[1], p.chosen);
        else
            std::snprintf(buf, sizeof buf, "pass %s %s %d ms %.7f - chosen %d\n", kPrec[p.pr], kPass[p.pass], p.L,
                          p.ms[0], p.chosen);
        o << buf;
    }
    for (const auto& f : r.fused) {
        std::snprintf(buf, sizeof buf, "yz %d %d ms %.7f %.7f chosen %d\n", f.ly, f.lz, f.ms_three, f.ms_fused,
                      f.chosen ? 1 : 0);
        o << buf;
    }
    for (const auto& p : r.planes) {
        std::snprintf(buf, sizeof buf, "xy %d ms %.7f %.7f chosen %d\n", p.L, p.ms_off, p.ms_on, p.chosen ? 1 : 0);
        o << buf;
    }
    return o.str();
}

}  // namespace tdda
