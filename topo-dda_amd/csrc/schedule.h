// Kernel schedule and its autotuner.
//
// The schedule records, per transform length, pass and precision, which split
// variant of the length-specialised kernels to run (with_split_pass), per
// (Ly, Lz) whether c64 transforms run the fused y/z pass, and per plane length
// L (Lx = Ly) whether they run the plane passes (x and y of a z-plane in one
// kernel per direction; these take precedence over the fused y/z pass).
// Entries are read at tile planning (plan_tiles, at System construction) and
// batch building, so they apply to systems and batches made afterwards.
// Absent entries mean variant 0, no fused y/z pass, and the plane passes on.
//
// The tuner times the candidates on the device for a set of box shapes and
// records the faster choice of each.  The schedule is plain text, one entry
// per line:
//   var <c64|c128> <xinv|y|z|xfwd> <L> <variant>
//   yz <Ly> <Lz> <0|1>
//   xy <L> <0|1>
// and is loaded at start-up from the file named by TDDA_SCHEDULE, if set.
#pragma once

#include <array>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "problem.h"

namespace tdda {

class Schedule {
   public:
    static Schedule& instance();
    int variant(int pr, int pass, int L);
    bool fused(int ly, int lz);
    bool plane(int L);
    void set_variant(int pr, int pass, int L, int v);
    void set_fused(int ly, int lz, bool on);
    void set_plane(int L, bool on);
    void clear();
    std::string text();
    // replaces every entry; a malformed line raises ERR_INVALID and leaves the
    // schedule unchanged
    void parse(const std::string& text);
    void load(const std::string& path);
    void save(const std::string& path);

   private:
    Schedule();
    std::mutex mu_;
    std::map<std::tuple<int, int, int>, int> var_;   // (pr, pass, L) -> variant
    std::map<std::pair<int, int>, int> fused_;       // (Ly, Lz) -> on
    std::map<int, int> plane_;                       // L -> on
};

struct TunePass {
    int pr, pass, L;
    double ms[2];   // summed over the boxes with this length; ms[1] < 0: no variant 1
    int chosen;
};
struct TuneFused {
    int ly, lz;
    double ms_three, ms_fused;   // whole c64 matvec, summed over the boxes of this class
    bool chosen;
};
struct TunePlane {
    int L;
    double ms_off, ms_on;   // whole c64 matvec, summed over the boxes of this length
    bool chosen;
};
struct TuneReport {
    std::vector<TunePass> passes;
    std::vector<TuneFused> fused;
    std::vector<TunePlane> planes;
};

// For `count` LDR problems of every box (random wavelengths and densities,
// seeded), time each pass of the c64 and c128 matvecs with every variant
// (reps batched launches per measurement), then the c64 matvec with the fused
// y/z pass against the three passes, both with the plane passes off, then per
// plane length the c64 matvec with the plane passes against the choices made
// so far; record the faster choices (a choice leaves the default only when
// the alternative is at least 3 % faster).
TuneReport tune(const std::vector<std::array<int, 3>>& boxes, int count, int reps, hipStream_t st);
std::string tune_report_text(const TuneReport& r);

}  // namespace tdda
