// engine.cpp -- bin/excent: flags, setup, the pass loop, results and report.
//
// Control flow, flags, defaults, messages and report lines follow the reference's
// main() (src/excent.cu:242-805, paths relative to the reference root) so scripts that
// drive or parse bin/excent keep working.  Device work goes through the C ABI of
// include/excent_hip.h; this file includes no HIP header.  Differences by design:
//   * resident chunks next to each other with the same list/range decision are pushed
//     in one launch (the arena holds them contiguously);
//   * the hot lists of every chunk are filled at setup (the reference fills them in
//     pass 0 only from chunks that stream, leaving lists of skipped chunks unfilled);
//   * the hot push reads its count on the device (no blocking count readback);
//   * tau calibration includes the PCIe copy of what it times;
//   * extension flags: --preload, --cache-mb MB, --push-kernel lb|group,
//     --ev-init outdeg|indeg, --stats-json FILE, --timing, --log FILE, --device N;
//   * a session API (excent_session_*) runs setup once and the pass loop many times
//     on the resident graph, for benchmarks.
#include <omp.h>
#include <sys/resource.h>
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "egr_graph.hpp"
#include "excent.h"
#include "excent_hip.h"
#include "planner.hpp"
#include "staging.hpp"

namespace excent {
namespace {

typedef std::chrono::steady_clock Clock;
double secs(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); }

const char* MNAME[EXC_MAXM] = {"pr", "katz", "ev"};

struct exit_error {
    int code;
    std::string msg;
};

[[noreturn]] void die(const std::string& msg, int code = 1) { throw exit_error{code, msg}; }

void ck(int rc, const char* what) {
    if (rc != EXC_OK) die(std::string("HIP layer error in ") + what + ": " + exc_last_error());
}

struct Options {
    std::string graph;
    double chunk_medges = 0, stream_ratio = 2.0, reserve_mb = 256;
    double target_chunks = 16, min_chunk_medges = 4, list_frac = 0.5;
    bool use_cache = true;
    std::string cache_mode = "auto", hot_when = "end", hot_order = "sweep", trace_file;
    bool stream_auto = true, recalibrate = false, dedup = true;
    std::string hot_rank;
    int group_opt = 0;
    bool count_pushes = false;
    std::string zero_copy = "auto";
    int max_passes = 100000;
    std::string dump, measure;
    double alpha = -1, beta = 1.0, lambda = -1, eps_pr = 0.01, eps_katz = -1, eps_ev = -1;
    // extensions
    bool preload = false;
    std::string drop_measures = "on";   // dense passes leave out measures with nothing pending
    int timing = 0;   // 1: HIP events around every kernel; 2: around the dense pull only
    double cache_mb = -1;
    std::string push_kernel = "lb", ev_init = "outdeg", stats_json, log;
    int device = 0;
    // batched personalised PageRank (--measure ppr)
    int ppr_k = 0;
    int ppr_col = -1;   // --measure pr with one batched column's teleport set (plain solver)
    std::string ppr_seeds, ppr_gen;
    std::string pull = "auto";   // dense passes by pull over a device transpose: auto | off | threshold
    double eps_ppr = -1;
    // vertex: stop when nothing is pending (the reference's rule); global: when each measure
    // either has nothing pending or sum |r_k| <= n eps_k; or a comma list of the measures
    // (pr, katz, ev) that take the global rule while the others keep the per-vertex one
    std::string stop = "vertex";
    int io_threads = 0;           // host threads for reads, gathers and checks (0: all)
    int hub_lds = 1024;           // destinations privatised in LDS by the hub-piece push (0: none)
    int burst = 0;                // passes per host synchronisation over a cached graph (0: automatic)
    std::string burst_ahead = "on";   // submit the next burst before reading the current one
};

const char* USAGE =
    "usage: %s graph.egr --measure pr|katz|ev|all [--alpha A] [--beta 1] [--lambda L] "
    "[--eps-pr 0.01] [--eps-katz E] [--eps-ev E] [--chunk-medges auto] [--target-chunks 16] "
    "[--min-chunk-medges 4] [--list-frac 0.5] [--stream-ratio auto|R] "
    "[--reserve-mb 256] [--cache chunks|hot|none] [--max-passes N] [--dump PREFIX]\n";

// src/excent.cu:269-299
Options parse(int argc, const char* const* argv) {
    Options o;
    if (argc < 2) {
        fprintf(stderr, USAGE, argc ? argv[0] : "excent");
        die("", 1);
    }
    o.graph = argv[1];
    for (int i = 2; i < argc; i++) {
        std::string a = argv[i];
        auto next = [&]() -> const char* {
            if (i + 1 >= argc) die("missing value for " + a);
            return argv[++i];
        };
        if (a == "--chunk-medges") o.chunk_medges = atof(next());
        else if (a == "--target-chunks") o.target_chunks = atof(next());
        else if (a == "--min-chunk-medges") o.min_chunk_medges = atof(next());
        else if (a == "--list-frac") o.list_frac = atof(next());
        else if (a == "--stream-ratio") { std::string v = next(); o.stream_auto = v == "auto"; if (!o.stream_auto) o.stream_ratio = atof(v.c_str()); }
        else if (a == "--dedup") { std::string v = next(); if (v != "on" && v != "off") die("--dedup must be on or off"); o.dedup = v == "on"; }
        else if (a == "--hot-rank") o.hot_rank = next();
        else if (a == "--recalibrate") o.recalibrate = true;
        else if (a == "--reserve-mb") o.reserve_mb = atof(next());
        else if (a == "--no-cache") o.use_cache = false;
        else if (a == "--cache") o.cache_mode = next();
        else if (a == "--hot-when") o.hot_when = next();
        else if (a == "--hot-order") o.hot_order = next();
        else if (a == "--trace-pending") o.trace_file = next();
        else if (a == "--group") o.group_opt = atoi(next());
        else if (a == "--count-pushes") o.count_pushes = true;
        else if (a == "--zero-copy") o.zero_copy = next();
        else if (a == "--max-passes") o.max_passes = atoi(next());
        else if (a == "--dump") o.dump = next();
        else if (a == "--measure") o.measure = next();
        else if (a == "--alpha") o.alpha = atof(next());
        else if (a == "--beta") o.beta = atof(next());
        else if (a == "--lambda") o.lambda = atof(next());
        else if (a == "--eps-pr") o.eps_pr = atof(next());
        else if (a == "--eps-katz") o.eps_katz = atof(next());
        else if (a == "--eps-ev") o.eps_ev = atof(next());
        // extensions (no reference counterpart)
        else if (a == "--preload") o.preload = true;
        else if (a == "--timing") o.timing = 1;
        else if (a == "--timing-pull") o.timing = 2;   // events around the dense pull only
        else if (a == "--cache-mb") o.cache_mb = atof(next());
        else if (a == "--push-kernel") o.push_kernel = next();
        else if (a == "--ev-init") o.ev_init = next();
        else if (a == "--stats-json") o.stats_json = next();
        else if (a == "--log") o.log = next();
        else if (a == "--device") o.device = atoi(next());
        else if (a == "--pull") o.pull = next();
        else if (a == "--ppr-k") o.ppr_k = atoi(next());
        else if (a == "--ppr-seeds") o.ppr_seeds = next();
        else if (a == "--ppr-gen") o.ppr_gen = next();
        else if (a == "--eps-ppr") o.eps_ppr = atof(next());
        else if (a == "--ppr-column") o.ppr_col = atoi(next());
        else if (a == "--stop") o.stop = next();
        else if (a == "--io-threads") o.io_threads = atoi(next());
        else if (a == "--hub-lds") o.hub_lds = atoi(next());
        else if (a == "--drop-measures") o.drop_measures = next();
        else if (a == "--burst") o.burst = atoi(next());
        else if (a == "--burst-ahead") o.burst_ahead = next();
        else die("unknown option " + a);
    }
    return o;
}

// Every column id must name a vertex: the push kernels add to r[id] with no bounds check.
void check_columns(const int32_t* a, size_t cnt, int n, const char* where, int threads) {
    long long bad = -1;
#pragma omp parallel for schedule(static) num_threads(threads) reduction(max : bad)
    for (long long i = 0; i < (long long)cnt; i++)
        if ((uint32_t)a[i] >= (uint32_t)n) bad = std::max(bad, i);
    if (bad >= 0) {
        char b[256];
        snprintf(b, sizeof b, "graph: column id %d out of range [0, %d) in %s", a[bad], n, where);
        die(b);
    }
}

// The tau cache of src/excent.cu:518-526, 576: ~/.cache/excent/<file name>.<n>.<m>.tau
// holding "%.6f %.4f" (tau, calibration seconds).
std::string tau_cache_path(const std::string& graph, int n, uint64_t m) {
    const char* home = getenv("HOME");
    std::string dir = std::string(home ? home : ".") + "/.cache";
    mkdir(dir.c_str(), 0755);
    dir += "/excent";
    mkdir(dir.c_str(), 0755);
    const size_t slash = graph.find_last_of('/');
    const std::string name = slash == std::string::npos ? graph : graph.substr(slash + 1);
    return dir + "/" + name + "." + std::to_string(n) + "." + std::to_string(m) + ".tau";
}

bool read_tau_cache(const std::string& path, double* tau, double* seconds) {
    FILE* f = fopen(path.c_str(), "r");
    if (!f) return false;
    const bool ok = fscanf(f, "%lf %lf", tau, seconds) == 2;
    fclose(f);
    return ok;
}

void write_tau_cache(const std::string& path, double tau, double seconds) {
    if (FILE* f = fopen(path.c_str(), "w")) {
        fprintf(f, "%.6f %.4f\n", tau, seconds);
        fclose(f);
    }
}

struct Faults { long maj, min; };
Faults faults() {
    struct rusage ru;
    getrusage(RUSAGE_SELF, &ru);
    return {ru.ru_majflt, ru.ru_minflt};
}

}  // namespace

class Engine {
public:
    explicit Engine(const Options& o, bool session = false) : o_(o), session_(session) {}
    ~Engine() {
        if (ctx_) exc_destroy(ctx_);
        if (out_ && own_out_) fclose(out_);
    }
    void setup();
    void run_passes();
    void reinit();
    void set_column(int j);
    void values(int col, int which, float* out) { ck(exc_read_state(ctx_, col, which, out, 0, (size_t)n_), "read state"); }
    int vertices() const { return n_; }
    void results();
    void report();
    void write_stats_json();
    void fill_stats(excent_run_stats* s);

private:
    Options o_;
    bool session_ = false, checkpoint_ = false;
    FILE* out_ = stdout;
    bool own_out_ = false;
    Clock::time_point T0_, T_open_, T_setup_, T_iter_, T_end_;
    std::unique_ptr<EgrGraph> g_;
    int n_ = 0, nc_ = 0, push_kernel_ = EXC_KERNEL_LB, cus_ = 0;
    int threads_ = 1;   // host threads for reads, gathers and checks
    uint64_t m_ = 0, cap_ = 0;
    uint64_t slot_edges_ = 0;           // staging slot capacity: edges, vertices
    int slot_verts_ = 0;
    char arch_[64] = {0};
    bool zc_ = false, use_cache_ = false, hot_mode_ = false;
    std::vector<EdgeChunk> chunks_;
    std::vector<int> kinds_;
    exc_measures M_{};
    void* ctx_ = nullptr;
    std::vector<char> resident_, loaded_;
    int n_resident_ = 0, nh_ = 0;
    uint64_t resident_edges_ = 0, hot_edges_ = 0, hot_deg_min_ = 0;
    std::vector<int> hlo_, hhi_;
    HotSet hs_;
    int32_t* h_edges_[2] = {nullptr, nullptr};
    int32_t* h_vids_[2] = {nullptr, nullptr};
    uint64_t* h_off_[2] = {nullptr, nullptr};
    double stream_ratio_ = 2.0, t_calib_ = 0;
    TauProbe tau_probe_;   // the calibration's measurements when it ran in this process
    // per run
    double t_io_ = 0, t_wait_ = 0, t_snap_ = 0;
    uint64_t bytes_stream_ = 0, bytes_gather_ = 0;
    int pass_ = 0, cur_ = 0;
    uint64_t n_list_ = 0, n_range_ = 0, n_launch_ = 0;
    long tot_maj_ = 0, tot_min_ = 0;
    uint64_t tot_pages_ = 0, tot_runs_ = 0;
    std::vector<double> sum_;
    std::vector<uint32_t> indeg_;
    double l1_[EXC_MAXM] = {0, 0, 0};   // FP64 sum of |r| per measure, last snapshot
    bool have_l1_ = false;
    int stop_reason_ = 0;               // 0 nothing pending, 1 global L1 bound, 2 pass limit
    std::vector<char> global_l1_;       // per measure slot: the global L1 rule may end it (--stop)
    int act_ = 1;                       // measure slots the passes update (exc_set_active_measures)
    bool dropped_global_ = false;       // a measure was frozen by the global rule with vertices above eps
    int burst_ = 1;                     // passes per burst (exc_run_passes); 1: every pass synchronises
    bool burst_ahead_ = true;           // --burst-ahead
    float linf_[EXC_MAXM] = {0, 0, 0};  // largest |r| per measure, last snapshot

    void init_state();
    void load_ppr_seeds(EgrGraph& g);
    void maybe_enable_pull();
    bool pull_tried_ = false, pull_on_ = false;
    double t_open_s_ = 0, t_setup_s_ = 0, t_transpose_s_ = 0;   // setup phases (outside the pass loop)
    double pull_thr_ = 1.0;   // dense pass (fold + pull) when demand >= pull_thr_ * m
    bool batched() const { return o_.measure == "ppr"; }
    int K_ = 0;
    std::vector<std::vector<int32_t>> seeds_;
    int ppr_col_ = -1;   // plain PageRank started from this column's teleport set
    std::vector<uint32_t> colpend_;
    int active_cols_ = 0;
    size_t budget_cap(size_t b) const {
        if (o_.cache_mb < 0) return b;
        return std::min(b, (size_t)(o_.cache_mb * (1 << 20)));
    }
};

void Engine::setup() {
    const Options& o = o_;
    T0_ = Clock::now();
    if (!o.log.empty()) {
        out_ = fopen(o.log.c_str(), "w");
        if (!out_) { out_ = stdout; die("cannot open --log file " + o.log); }
        own_out_ = true;
    }
    FILE* out = out_;

    // src/excent.cu:301-316
    std::string cache_mode = o.use_cache ? o.cache_mode : "none";
    if (cache_mode != "chunks" && cache_mode != "hot" && cache_mode != "none" && cache_mode != "auto")
        die("--cache must be auto, chunks, hot or none");
    use_cache_ = cache_mode == "chunks";
    hot_mode_ = cache_mode == "hot";
    if (o.measure == "pr") kinds_ = {EXC_PR};
    else if (o.measure == "katz") kinds_ = {EXC_KATZ};
    else if (o.measure == "ev") kinds_ = {EXC_EV};
    else if (o.measure == "all") kinds_ = {EXC_PR, EXC_KATZ, EXC_EV};
    else if (o.measure == "ppr") kinds_ = {EXC_PR};
    else die("--measure must be pr, katz, ev or all");
    if (o.measure == "ppr") {
        if (o.eps_ppr <= 0) die("ppr needs --eps-ppr");
        if (o.ppr_seeds.empty() == o.ppr_gen.empty()) die("ppr needs one of --ppr-seeds FILE or --ppr-gen K:S:TOP:SEED");
        if (!o.ppr_seeds.empty() && (o.ppr_k < 1 || o.ppr_k > 64)) die("--ppr-k must be 1..64");
    }
    if (o.ppr_col >= 0) {
        if (o.measure != "pr") die("--ppr-column needs --measure pr");
        if (o.ppr_seeds.empty() == o.ppr_gen.empty()) die("--ppr-column needs one of --ppr-seeds FILE or --ppr-gen K:S:TOP:SEED");
        if (!o.ppr_seeds.empty() && (o.ppr_k < 1 || o.ppr_k > 64)) die("--ppr-k must be 1..64");
    }
    for (int k : kinds_) {
        if (k == EXC_KATZ && (o.alpha <= 0 || o.eps_katz <= 0)) die("katz needs --alpha and --eps-katz");
        if (k == EXC_EV && (o.lambda <= 0 || o.eps_ev <= 0)) die("ev needs --lambda and --eps-ev");
    }
    if (o.push_kernel == "lb") push_kernel_ = EXC_KERNEL_LB;
    else if (o.push_kernel == "group") push_kernel_ = EXC_KERNEL_GROUP;
    else if (o.push_kernel == "tune") push_kernel_ = EXC_KERNEL_TUNE;
    else die("--push-kernel must be lb, group or tune");
    if (o.ev_init != "outdeg" && o.ev_init != "indeg") die("--ev-init must be outdeg or indeg");
    if (o.pull != "auto" && o.pull != "off" && atof(o.pull.c_str()) <= 0) die("--pull must be auto, off or a positive demand fraction");
    if (o.zero_copy != "auto" && o.zero_copy != "on" && o.zero_copy != "off") die("--zero-copy must be auto, on or off");
    if (o.group_opt < 0 || o.group_opt > 64) die("--group must be 0 (automatic) or 1..64");
    if (o.hot_when != "end" && o.hot_when != "start" && o.hot_when != "both") die("--hot-when must be end, start or both");
    if (o.hot_order != "sweep" && o.hot_order != "ranges") die("--hot-order must be sweep or ranges");
    if (o.hub_lds < 0 || o.hub_lds > 1024) die("--hub-lds must be 0..1024");
    if (o.drop_measures != "on" && o.drop_measures != "off") die("--drop-measures must be on or off");
    if (o.burst_ahead != "on" && o.burst_ahead != "off") die("--burst-ahead must be on or off");
    if (o.burst < 0 || o.burst > EXC_BURST_MAX) die("--burst must be 0 (automatic) or 1.." + std::to_string(EXC_BURST_MAX));
    global_l1_.assign(kinds_.size(), 0);
    if (o.stop == "global") {
        global_l1_.assign(kinds_.size(), 1);
    } else if (o.stop != "vertex") {
        size_t at = 0;
        while (at <= o.stop.size()) {
            const size_t comma = std::min(o.stop.find(',', at), o.stop.size());
            const std::string name = o.stop.substr(at, comma - at);
            bool found = false;
            for (size_t i = 0; i < kinds_.size(); i++)
                if (name == MNAME[kinds_[i]]) { global_l1_[i] = 1; found = true; }
            if (!found) die("--stop must be vertex, global or a list of this run's measures (pr, katz, ev)");
            at = comma + 1;
        }
    }
    if (o.stop != "vertex" && o.measure == "ppr") die("--stop global needs plain measures (pr, katz, ev, all)");

    if (o.io_threads < 0) die("--io-threads must be 0 (every host thread) or a thread count");
    try {
        g_.reset(new EgrGraph(o.graph, o.io_threads, /*offsets_in_ram=*/false));
    } catch (const egr_error& e) {
        die(e.what());
    }
    threads_ = g_->io_threads();
    EgrGraph& g = *g_;
    n_ = g.n;
    m_ = g.m;
    const int n = n_;
    const uint64_t m = m_;
    cap_ = chunk_cap(m, o.chunk_medges, o.target_chunks, o.min_chunk_medges);
    chunks_ = g.plan_chunks(cap_);
    nc_ = (int)chunks_.size();
    const int nc = nc_;
    uint64_t max_chunk_edges = 0;
    int max_chunk_verts = 0;
    for (auto& c : chunks_) { max_chunk_edges = std::max(max_chunk_edges, c.edges()); max_chunk_verts = std::max(max_chunk_verts, c.v1 - c.v0); }
    T_open_ = Clock::now();
    fprintf(out, "graph: n=%d m=%llu chunks=%d (cap %llu, max %llu edges, %d vertices)  open %.2fs\n",
            n, (unsigned long long)m, nc, (unsigned long long)cap_, (unsigned long long)max_chunk_edges, max_chunk_verts,
            secs(T0_, T_open_));

    int integrated = 0;
    ck(exc_device_info(o.device, &cus_, &integrated, arch_, sizeof arch_), "device info");
    zc_ = o.zero_copy == "on" || (o.zero_copy == "auto" && integrated);

    M_.nm = (int)kinds_.size();
    for (int i = 0; i < M_.nm; i++) {
        const int k = kinds_[i];
        M_.kind[i] = k;
        M_.eps[i] = (float)(batched() || (o.ppr_col >= 0 && o.eps_ppr > 0) ? o.eps_ppr
                            : k == EXC_PR ? o.eps_pr : k == EXC_KATZ ? o.eps_katz : o.eps_ev);
        M_.scale[i] = (float)(k == EXC_PR ? (double)0.85f : k == EXC_KATZ ? o.alpha : 1.0 / o.lambda);
        M_.divdeg[i] = k == EXC_PR;
    }
    std::vector<int> cv0(nc);
    for (int c = 0; c < nc; c++) cv0[c] = chunks_[c].v0;
    exc_config cfg{};
    cfg.device = o.device; cfg.n = n; cfg.m = m; cfg.nc = nc; cfg.cv0 = cv0.data();
    cfg.max_chunk_edges = max_chunk_edges; cfg.max_chunk_verts = max_chunk_verts;
    slot_edges_ = max_chunk_edges;
    slot_verts_ = max_chunk_verts;
    cfg.trace = !o.trace_file.empty(); cfg.zero_copy = zc_; cfg.push_kernel = push_kernel_; cfg.timing = o.timing;
    if (batched()) {
        load_ppr_seeds(g);
        cfg.batch_k = K_;
    } else if (o.ppr_col >= 0) {
        load_ppr_seeds(g);
        if (o.ppr_col >= K_) die("--ppr-column must be below the number of columns");
        ppr_col_ = o.ppr_col;
    }
    ck(exc_create(&cfg, &M_, &ctx_), "create");
    void* ctx = ctx_;
    {   // offsets in 16M-entry slices (src/excent.cu:360-369)
        const size_t B = (size_t)1 << 24;
        std::vector<uint64_t> stage(std::min(B, (size_t)n + 1));
        for (size_t v0 = 0; v0 < (size_t)n + 1; v0 += B) {
            const size_t cnt = std::min(B, (size_t)n + 1 - v0);
            g.read_offsets(v0, cnt, stage.data());
            ck(exc_upload_offsets(ctx, stage.data(), v0, cnt), "offsets upload");
        }
    }
    if (o.ev_init == "indeg" && std::find(kinds_.begin(), kinds_.end(), (int)EXC_EV) != kinds_.end()) {
        indeg_.assign(n, 0);
        const int32_t* nl = g.nlist();
#pragma omp parallel for schedule(static) num_threads(threads_)
        for (long long e = 0; e < (long long)m; e++) __atomic_fetch_add(&indeg_[nl[e]], 1u, __ATOMIC_RELAXED);
    }
    init_state();

    resident_.assign(nc, 0);
    loaded_.assign(nc, 0);
    const size_t reserve = (size_t)(o.reserve_mb * (1 << 20));
    if (cache_mode == "auto") {   // src/excent.cu:402-409
        size_t freeB, totalB;
        ck(exc_mem_info(o.device, &freeB, &totalB), "mem info");
        const size_t avail = budget_cap(freeB > reserve ? freeB - reserve : 0);
        const bool fits = freeB > reserve && m * sizeof(int32_t) <= avail;
        use_cache_ = fits;
        hot_mode_ = !fits;
        fprintf(out, "cache: auto -> %s\n", fits ? "chunks (the whole graph fits)" : "high-degree lists");
    }
    if (use_cache_) {   // src/excent.cu:410-421: resident chunks are a prefix in file order
        size_t freeB, totalB;
        ck(exc_mem_info(o.device, &freeB, &totalB), "mem info");
        size_t avail = budget_cap(freeB > reserve ? freeB - reserve : 0);
        for (int c = 0; c < nc; c++) {
            const size_t need = chunks_[c].edges() * sizeof(int32_t);
            if (need > avail) break;
            avail -= need;
            n_resident_++;
        }
        while (n_resident_ > 0) {
            const int rc = exc_cache_chunks(ctx, chunks_[n_resident_ - 1].e1);
            if (rc == EXC_OK) break;
            if (rc != EXC_ENOMEM) ck(rc, "edge cache");
            n_resident_--;   // allocation failure: keep a shorter prefix
        }
        for (int c = 0; c < n_resident_; c++) { resident_[c] = 1; resident_edges_ += chunks_[c].edges(); }
    }

    hlo_.assign(nc, 0);
    hhi_.assign(nc, 0);
    if (hot_mode_) {   // src/excent.cu:428-513
        size_t freeB, totalB;
        ck(exc_mem_info(o.device, &freeB, &totalB), "mem info");
        const size_t bits = (((size_t)n + 31) / 32) * sizeof(uint32_t);
        const size_t budget = budget_cap(freeB > reserve + bits ? freeB - reserve - bits : 0);
        std::vector<uint16_t> score;
        if (!o.hot_rank.empty()) {
            score.resize(n);
            FILE* f = fopen(o.hot_rank.c_str(), "rb");
            if (!f || fread(score.data(), sizeof(uint16_t), n, f) != (size_t)n) {
                if (f) fclose(f);
                char b[512];
                snprintf(b, sizeof b, "--hot-rank: cannot read %d entries from %s", n, o.hot_rank.c_str());
                die(b);
            }
            fclose(f);
        }
        hs_ = hot_select(n, g.offsets(), budget, score.empty() ? nullptr : score.data(), threads_);
        nh_ = (int)hs_.ids.size();
        hot_edges_ = hs_.off.back();
        hot_deg_min_ = hs_.deg_min;
        const int rc = nh_ > 0 ? exc_cache_hot(ctx, nh_, hs_.ids.data(), hs_.off.data(), hs_.bitmask.data()) : EXC_ENOMEM;
        if (rc != EXC_OK) {
            if (rc != EXC_ENOMEM) ck(rc, "hot cache");
            fprintf(stderr, "degree cache: could not allocate %d lists (%.2f GB); running without a cache\n", nh_, hot_edges_ * 4.0 / 1e9);
            nh_ = 0;
            hot_edges_ = 0;
        } else {
            for (int c = 0; c < nc; c++) {
                hlo_[c] = (int)(std::lower_bound(hs_.ids.begin(), hs_.ids.end(), chunks_[c].v0) - hs_.ids.begin());
                hhi_[c] = (int)(std::lower_bound(hs_.ids.begin(), hs_.ids.end(), chunks_[c].v1) - hs_.ids.begin());
            }
        }
    }

    for (int s = 0; s < 2; s++) ck(exc_slot_buffers(ctx, s, &h_edges_[s], &h_vids_[s], &h_off_[s]), "slot buffers");

    if (nh_ > 0) {   // every chunk's hot lists are filled here, before the passes
        for (int c = 0; c < nc; c++) {
            if (hhi_[c] <= hlo_[c]) continue;
            const EdgeChunk& ch = chunks_[c];
            g.read_edges(ch.e0, ch.e1, h_edges_[0]);
            check_columns(h_edges_[0], ch.edges(), n, "a hot chunk", threads_);
            const uint64_t o0 = hs_.off[hlo_[c]], o1 = hs_.off[hhi_[c]];
            int32_t* dst = h_edges_[1];
            const int32_t* src = h_edges_[0];
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads_)
            for (int j = hlo_[c]; j < hhi_[c]; j++)
                memcpy(dst + (hs_.off[j] - o0), src + (g.offset(hs_.ids[j]) - ch.e0), (hs_.off[j + 1] - hs_.off[j]) * sizeof(int32_t));
            ck(exc_cache_hot_upload(ctx, dst, o0, o1 - o0), "hot fill");
        }
    }
    if (o.preload && n_resident_ > 0) {   // extension: fill the chunk cache before the passes
        for (int c = 0; c < n_resident_; c++) {
            const EdgeChunk& ch = chunks_[c];
            g.read_edges(ch.e0, ch.e1, h_edges_[0]);
            check_columns(h_edges_[0], ch.edges(), n, "a preloaded chunk", threads_);
            ck(exc_cache_upload(ctx, h_edges_[0], ch.e0, ch.edges()), "preload");
            loaded_[c] = 1;
        }
    }

    // tau: given, cached for this graph on this machine, or measured (src/excent.cu:515-579
    // defines the rule, the cache file and the report lines; the measurement is staging.cpp's)
    stream_ratio_ = o.stream_ratio;
    if (!o.stream_auto) {
        fprintf(out, "fetch rule: tau = %.3f (given)\n", stream_ratio_);
    } else {
        const std::string tau_file = tau_cache_path(o.graph, n, m);
        if (!o.recalibrate && read_tau_cache(tau_file, &stream_ratio_, &t_calib_)) {
            fprintf(out, "fetch rule: tau = %.3f (measured earlier on this machine in %.2fs, %s)\n", stream_ratio_, t_calib_,
                    tau_file.c_str());
        } else {
            // the largest chunk: the one whose fetch costs the most
            const EdgeChunk* ch = &chunks_[0];
            for (const EdgeChunk& c : chunks_) if (c.edges() > ch->edges()) ch = &c;
            tau_probe_ = calibrate_tau(g, *ch, h_edges_[0], h_vids_[0], h_off_[0], o.dedup, threads_, [&](uint64_t bytes) {
                double t = 0;
                ck(exc_slot_copy_probe(ctx, 0, bytes, &t), "calibration copy");
                return t;
            });
            const TauProbe& p = tau_probe_;
            stream_ratio_ = p.tau;
            t_calib_ = p.seconds;
            fprintf(out, "fetch rule: tau = %.3f (stream %.3fs; gather %.3fs / %.3fs / %.3fs at demand %.3f / %.3f / %.3f)  calibrate %.2fs\n",
                    p.tau, p.t_stream, p.t_gather[0], p.t_gather[1], p.t_gather[2], p.f_got[0], p.f_got[1], p.f_got[2], p.seconds);
            fprintf(out, "fetch rule: storage %.2f GB/s (cold chunk read), link %.2f GB/s (pinned host to GPU), %d host threads\n",
                    p.storage_seq_gbs, p.link_gbs, threads_);
            write_tau_cache(tau_file, stream_ratio_, t_calib_);
        }
    }
    if (o.hub_lds > 0 && !batched()) {   // destinations with the most in-edges, from 1/16 of the columns
        const std::vector<int32_t> hubs = top_in_degree(g, o.hub_lds, 1.0 / 16, threads_);
        ck(exc_set_hub_destinations(ctx, hubs.data(), (int)hubs.size()), "hub destinations");
    }
    maybe_enable_pull();
    // bursts pay where a pass is short next to a host round trip: graphs of up to 2^25 edges
    burst_ = o.burst ? o.burst : (m_ <= (1ull << 25) ? 16 : 1);
    burst_ahead_ = o.burst_ahead == "on";
    ck(exc_pass_sync(ctx, nullptr), "setup sync");
    ck(exc_reset_stats(ctx), "reset stats");
    checkpoint_ = session_ && exc_state_checkpoint(ctx, 0) == EXC_OK;
    T_setup_ = Clock::now();
    t_open_s_ = secs(T0_, T_open_);
    t_setup_s_ = secs(T_open_, T_setup_);
    fprintf(out, "staging: %s\n", zc_ ? "zero copy (pinned, GPU-mapped buffers read in place)" : "pinned host buffers copied to the GPU");
    if (hot_mode_ && !o.hot_rank.empty())
        fprintf(out, "GPU cache ranked by passes pending (%s): %d vertices (%.1f%% of edges, %.2f GB)  setup %.2fs\n",
                o.hot_rank.c_str(), nh_, m ? 100.0 * hot_edges_ / m : 0.0, hot_edges_ * 4.0 / 1e9, secs(T_open_, T_setup_));
    else if (hot_mode_)
        fprintf(out, "GPU degree cache: %d vertices with degree >= %llu (%.1f%% of edges, %.2f GB)  setup %.2fs\n",
                nh_, (unsigned long long)hot_deg_min_, m ? 100.0 * hot_edges_ / m : 0.0, hot_edges_ * 4.0 / 1e9, secs(T_open_, T_setup_));
    else
        fprintf(out, "GPU edge cache: %d/%d chunks resident (%.1f%% of edges)  setup %.2fs\n",
                n_resident_, nc, m ? 100.0 * resident_edges_ / m : 0.0, secs(T_open_, T_setup_));
}

// Teleport sets of the batched columns: from a file of K*S int32 vertex ids, or drawn
// with --ppr-gen K:S:TOP:SEED -- S distinct seeds per column, TOP of them uniform over the
// top 1% of vertices by out-degree and the rest uniform over all vertices (SplitMix64).
void Engine::load_ppr_seeds(EgrGraph& g) {
    const Options& o = o_;
    if (!o.ppr_seeds.empty()) {
        K_ = o.ppr_k;
        FILE* f = fopen(o.ppr_seeds.c_str(), "rb");
        if (!f) die("cannot read --ppr-seeds " + o.ppr_seeds);
        std::vector<int32_t> all;
        int32_t b[4096];
        size_t r;
        while ((r = fread(b, 4, 4096, f)) > 0) all.insert(all.end(), b, b + r);
        fclose(f);
        if (all.empty() || all.size() % K_) die("--ppr-seeds: the file must hold K * S int32 ids");
        const size_t S = all.size() / K_;
        seeds_.assign(K_, {});
        for (int j = 0; j < K_; j++) seeds_[j].assign(all.begin() + j * S, all.begin() + (j + 1) * S);
    } else {
        int Kg = 0, S = 0, top = 0;
        unsigned long long sd = 0;
        if (sscanf(o.ppr_gen.c_str(), "%d:%d:%d:%llu", &Kg, &S, &top, &sd) != 4 || Kg < 1 || Kg > 64 || S < 1 ||
            top < 0 || top > S || S > n_)
            die("--ppr-gen must be K:S:TOP:SEED with 1 <= K <= 64, TOP <= S <= n");
        K_ = Kg;
        std::vector<int> order(n_);
        for (int v = 0; v < n_; v++) order[v] = v;
        const uint64_t* off = g.offsets();
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return off[a + 1] - off[a] > off[b + 1] - off[b]; });
        const int ntop = std::max(top > 0 ? 1 : 0, n_ / 100);
        uint64_t state = sd * 0x9E3779B97F4A7C15ull + 12345;
        auto next = [&]() {
            uint64_t z = (state += 0x9E3779B97F4A7C15ull);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            return z ^ (z >> 31);
        };
        seeds_.assign(K_, {});
        for (int j = 0; j < K_; j++) {
            std::vector<int32_t>& sj = seeds_[j];
            while ((int)sj.size() < S) {
                const bool from_top = (int)sj.size() < top;
                const int v = from_top ? order[next() % (uint64_t)ntop] : (int)(next() % (uint64_t)n_);
                if (std::find(sj.begin(), sj.end(), v) == sj.end()) sj.push_back(v);
            }
        }
    }
    for (auto& sj : seeds_)
        for (int v : sj)
            if (v < 0 || v >= n_) die("ppr seed out of range");
}

// Once the whole graph sits in the chunk cache, dense passes can run as fold + pull over a
// transpose built on the device (extension; --pull off keeps the push only).
void Engine::maybe_enable_pull() {
    if (pull_tried_ || o_.pull == "off" || !use_cache_ || n_resident_ != nc_) return;
    for (int c = 0; c < nc_; c++) if (!loaded_[c]) return;
    pull_tried_ = true;
    // auto: 5 % of the edges (batched K = 64 at R-MAT scale 26: 4.19 s at 4 % and 4.18 s at
    // 8 %, 4.26 s at 2 %, 4.30 s at 1 %)
    const double thr = o_.pull == "auto" ? 0.05 : atof(o_.pull.c_str());
    const auto t0 = Clock::now();
    const int rc = exc_enable_pull(ctx_, thr);
    t_transpose_s_ = secs(t0, Clock::now());
    if (rc == EXC_OK) { pull_on_ = true; pull_thr_ = thr; }
    else if (rc != EXC_ENOMEM) ck(rc, "pull setup");
}

// k_init for every measure; the eigenvector's residual uses in-degrees with --ev-init indeg.
// Batched columns start at x = r = 0 with r = (1 - d) * n / S at each of a column's S seeds
// (teleport mass n per column, so values keep the unnormalised scale of mean about 1).
void Engine::init_state() {
    if (batched()) {
        ck(exc_init_measure(ctx_, 0, 0.f, 0.f), "init");
        for (int j = 0; j < K_; j++) {
            const float mass = (float)((1.0 - (double)0.85f) * (double)n_ / (double)seeds_[j].size());
            ck(exc_batch_seed(ctx_, j, seeds_[j].data(), (int)seeds_[j].size(), mass), "ppr seed");
        }
        return;
    }
    if (ppr_col_ >= 0) {   // one column of the batched problem, on the plain solver
        const std::vector<int32_t>& sj = seeds_[ppr_col_];
        const float mass = (float)((1.0 - (double)0.85f) * (double)n_ / (double)sj.size());
        std::vector<float> x(n_, 0.f), r(n_, 0.f);
        for (int v : sj) r[v] = mass;
        ck(exc_set_state(ctx_, 0, x.data(), r.data()), "ppr column init");
        return;
    }
    for (int i = 0; i < M_.nm; i++) ck(exc_init_measure(ctx_, i, (float)o_.beta, (float)o_.lambda), "init");
    if (indeg_.empty()) return;
    // r = indeg/lambda - 1 with x = 1: the invariant x + r = W x on a directed graph (the
    // reference's out-degree form assumes a symmetric graph).
    for (int i = 0; i < M_.nm; i++) {
        if (kinds_[i] != EXC_EV) continue;
        std::vector<float> r(n_), x(n_, 1.f);
        const double lam = (double)(float)o_.lambda;
#pragma omp parallel for schedule(static) num_threads(threads_)
        for (int v = 0; v < n_; v++) r[v] = (float)((double)indeg_[v] / lam - 1.0);
        ck(exc_set_state(ctx_, i, x.data(), r.data()), "ev init");
    }
}

// Switches a --ppr-column session to another column: fresh state, statistics cleared.
void Engine::set_column(int j) {
    if (ppr_col_ < 0) die("set_column: the session was not opened with --ppr-column");
    if (j < 0 || j >= K_) die("set_column: column out of range");
    ppr_col_ = j;
    init_state();
    ck(exc_pass_sync(ctx_, nullptr), "column sync");
    ck(exc_reset_stats(ctx_), "reset stats");
}

void Engine::reinit() {
    // restore the initial state saved at setup (a device copy); re-initialise if it
    // could not be saved
    if (!checkpoint_ || exc_state_checkpoint(ctx_, 1) != EXC_OK) init_state();
    ck(exc_pass_sync(ctx_, nullptr), "reinit sync");
    ck(exc_reset_stats(ctx_), "reset stats");
}

// ---- passes (src/excent.cu:593-760)
void Engine::run_passes() {
    const Options& o = o_;
    EgrGraph& g = *g_;
    void* ctx = ctx_;
    FILE* out = out_;
    const int n = n_, nc = nc_;
    const uint64_t m = m_;
    t_io_ = t_wait_ = t_snap_ = 0;
    bytes_stream_ = bytes_gather_ = 0;
    pass_ = 0;
    n_list_ = n_range_ = n_launch_ = 0;
    tot_maj_ = tot_min_ = 0;
    tot_pages_ = tot_runs_ = 0;
    stop_reason_ = 2;
    have_l1_ = false;
    bool stop_global = false;
    for (char g1 : global_l1_) stop_global = stop_global || g1;
    if (pull_on_) ck(exc_fold_ahead(ctx, 1), "fold ahead");
    act_ = M_.nm;   // a run starts from a replaced state: every measure active
    dropped_global_ = false;
    bool drop_wait = false;        // a measure waits to be dropped until a pass has not folded
    bool fold_ahead_off = false;   // set once a folded state met the global bound
    const Clock::time_point T_begin = Clock::now();
    const size_t words = ((size_t)n + 31) / 32;
    std::vector<uint64_t> h_cnt(2 * (size_t)nc);
    std::vector<uint32_t> h_bitmap(words);
    std::vector<uint64_t> chunk_edges(nc);
    for (int c = 0; c < nc; c++) chunk_edges[c] = chunks_[c].edges();
    uint64_t h_hotpend = 0;
    const int gcap = 32;
    const int G_hot = o.group_opt ? o.group_opt : pick_group(nh_ ? (double)hot_edges_ / nh_ : 32.0, gcap);
    std::vector<std::pair<int, int>> hot_ranges;
    if (o.hot_order == "ranges") { for (int c = 0; c < nc; c++) if (hhi_[c] > hlo_[c]) hot_ranges.push_back({hlo_[c], hhi_[c]}); }
    else if (nh_) hot_ranges.push_back({0, nh_});
    auto launch_hot = [&](int after_slots) {
        for (auto [lo, hi] : hot_ranges) ck(exc_push_hot(ctx, lo, hi, G_hot, after_slots), "hot push");
    };
    auto group_of = [&](uint64_t k, uint64_t e) {
        if (o.group_opt) return o.group_opt;
        return pick_group(k ? (double)e / k : 1.0, gcap);
    };
    std::vector<int8_t> mode(nc);
    // the report line of one pass, from the snapshot it started from (src/excent.cu:754-759)
    auto pass_line = [&](int pass, uint64_t pend_v, uint64_t demand, int nres, int nstream, int ngather, uint64_t hotpend,
                         long maj, long mn, uint64_t pages, uint64_t runs, bool have, const double* l1, const float* linf) {
        fprintf(out, "pass %3d: pending %llu v, demand ratio %.4f | resident %d stream %d gather %d | cached high-degree pending %llu"
                     " | faults major %ld minor %ld | gather pages %llu runs %llu",
                pass, (unsigned long long)pend_v, m ? (double)demand / m : 0.0, nres, nstream, ngather,
                (unsigned long long)hotpend, maj, mn, (unsigned long long)pages, (unsigned long long)runs);
        if (batched()) fprintf(out, " | active columns %d", active_cols_);
        if (have) {
            fprintf(out, " | residual l1");
            for (int i = 0; i < M_.nm; i++) fprintf(out, " %s %.4e", MNAME[kinds_[i]], l1[i]);
            fprintf(out, " | max |r|");
            for (int i = 0; i < M_.nm; i++) fprintf(out, " %s %.4e", MNAME[kinds_[i]], (double)linf[i]);
        }
        fprintf(out, "\n");
        fflush(out);   // a long run shows its progress pass by pass
    };
    bool whole_cached = !hot_mode_ && n_resident_ == nc;
    std::vector<uint64_t> b_cnt;
    std::vector<double> b_l1;
    std::vector<float> b_linf;
    for (; pass_ < o.max_passes; pass_++) {
        auto a = Clock::now();
        const Faults f0 = faults();
        uint64_t pass_pages = 0, pass_runs = 0;
        const int hot_skip = (hot_mode_ && nh_ > 0 && pass_ > 0) ? 1 : 0;
        ck(exc_snapshot(ctx, hot_skip, h_cnt.data(), &h_hotpend), "snapshot");
        if (batched()) {   // converged columns drop out of the pushes
            colpend_.resize(K_);
            ck(exc_batch_colpend(ctx, colpend_.data()), "column counts");
            uint64_t mask = 0;
            active_cols_ = 0;
            for (int j = 0; j < K_; j++) if (colpend_[j]) { mask |= 1ull << j; active_cols_++; }
            ck(exc_batch_set_active(ctx, mask), "active columns");
        }
        const PlanOut po = fetch_plan(nc, h_cnt.data(), resident_, loaded_, stream_ratio_, chunk_edges, hot_mode_, pass_, hlo_, hhi_, mode.data());
        // FP64 global residual of the snapshot state (plain measures)
        have_l1_ = !batched() && exc_residual_l1(ctx, l1_) == EXC_OK && exc_residual_max(ctx, linf_) == EXC_OK;
        drop_wait = false;
        if (have_l1_ && M_.nm > 1 && o.drop_measures == "on") {
            // A measure is done once nothing is pending in it, or, where the global rule may end
            // it (--stop), once its FP64 sum |r| <= n eps; the passes then leave out the done
            // measures after the last one still going (they are frozen: never pushed again, so
            // their state and sums stay as the rule found them).
            int act = 0;
            for (int i = 0; i < M_.nm; i++) {
                const bool done = linf_[i] <= M_.eps[i] || (global_l1_[i] && l1_[i] <= (double)n * (double)M_.eps[i]);
                if (!done) act = i + 1;
            }
            if (act >= 1 && act < act_) {
                if (exc_set_active_measures(ctx, act) == EXC_OK) {
                    for (int i = act; i < act_; i++) dropped_global_ = dropped_global_ || linf_[i] > M_.eps[i];
                    act_ = act;
                } else if (exc_dense_required(ctx) == 1) {
                    drop_wait = true;   // the last pull folded a pass that holds their deltas: no fold this pass
                } else {
                    die(std::string("HIP layer error in active measures: ") + exc_last_error());
                }
            }
        }
        bool global_met = false;
        if (stop_global && have_l1_) {
            // every measure done: nothing pending for it, or (global rule) its sum within n eps
            global_met = true;
            for (int i = 0; i < M_.nm; i++)
                global_met = global_met && (linf_[i] <= M_.eps[i] || (global_l1_[i] && l1_[i] <= (double)n * (double)M_.eps[i]));
            if (global_met && exc_dense_required(ctx) == 1) {
                // the last dense pass folded this state: its deltas are still in flight, so the
                // state cannot be returned; deliver them with no further fold and test again
                global_met = false;
                fold_ahead_off = true;
            }
        }
        if (po.need_bitmap && !global_met) ck(exc_bitmap_to_host(ctx, h_bitmap.data()), "bitmap readback");
        t_snap_ += secs(a, Clock::now());
        if (po.pend_v == 0 && h_hotpend == 0) { stop_reason_ = dropped_global_ ? 1 : 0; break; }
        if (global_met) { stop_reason_ = 1; break; }
        // a dense pass folds the pass after it only when the loop will go on from its state
        if (pull_on_) ck(exc_fold_ahead(ctx, !fold_ahead_off && !drop_wait && pass_ + 1 < o.max_passes), "fold ahead");

        // A burst: the next burst_ passes over the whole cached graph enqueued together with one
        // host synchronisation (exc_burst_submit / exc_burst_collect).  The per-vertex rule ends
        // the run at the first snapshot with nothing pending; the passes after it changed
        // nothing.  With --burst-ahead on the next burst (same kind) is submitted before the
        // current one is read, so the GPU runs it while the host reads, logs and decides; it
        // goes on so while the kind stays the one the demand picks and no measure finishes.
        if (whole_cached) for (int c = 0; c < nc && whole_cached; c++) whole_cached = loaded_[c] != 0;
        if (burst_ > 1 && whole_cached && o.trace_file.empty() && !stop_global && !batched() &&
            pass_ + burst_ < o.max_passes) {
            auto want_dense = [&](const uint64_t* cn) {
                uint64_t demand = 0;
                for (int c = 0; c < nc; c++) demand += cn[2 * c + 1];
                return pull_on_ && ((double)demand >= pull_thr_ * (double)m || exc_dense_required(ctx) == 1);
            };
            const bool dense = want_dense(h_cnt.data());
            const int B = burst_, nm = M_.nm;
            b_cnt.resize((size_t)B * 2 * nc);
            b_l1.resize((size_t)B * nm);
            b_linf.resize((size_t)B * nm);
            // the snapshot each reported pass started from: before the burst, the loop's own
            std::vector<uint64_t> p_cnt(h_cnt);
            std::vector<double> p_l1(l1_, l1_ + nm);
            std::vector<float> p_linf(linf_, linf_ + nm);
            // the report lines of the burst just collected (passes from pass_ on); false once a
            // pass ended with nothing pending (pass_ is then the run's pass count)
            auto report = [&](long maj, long mn) {
                const int first = pass_;
                for (int i = 0; i < B; i++) {
                    const uint64_t* cn = i == 0 ? p_cnt.data() : &b_cnt[(size_t)(i - 1) * 2 * nc];
                    uint64_t pv = 0, dm = 0;
                    int nr = 0;
                    for (int c = 0; c < nc; c++) { pv += cn[2 * c]; dm += cn[2 * c + 1]; nr += cn[2 * c] ? 1 : 0; }
                    pass_line(first + i, pv, dm, nr, 0, 0, 0, i == 0 ? maj : 0, i == 0 ? mn : 0, 0, 0, true,
                              i == 0 ? p_l1.data() : &b_l1[(size_t)(i - 1) * nm],
                              i == 0 ? p_linf.data() : &b_linf[(size_t)(i - 1) * nm]);
                    n_launch_++;
                    if (dense) n_range_ += nr; else n_list_ += nr;
                    uint64_t after = 0;
                    for (int c = 0; c < nc; c++) after += b_cnt[(size_t)i * 2 * nc + 2 * c];
                    if (after == 0) { pass_ = first + i + 1; return false; }
                }
                pass_ = first + B;
                std::copy(&b_cnt[(size_t)(B - 1) * 2 * nc], &b_cnt[(size_t)B * 2 * nc], p_cnt.begin());
                std::copy(&b_l1[(size_t)(B - 1) * nm], &b_l1[(size_t)B * nm], p_l1.begin());
                std::copy(&b_linf[(size_t)(B - 1) * nm], &b_linf[(size_t)B * nm], p_linf.begin());
                return true;
            };
            ck(exc_burst_submit(ctx, B, dense ? 1 : 0), "burst");
            int inflight = 1;
            bool converged = false;
            Faults fb = f0;
            for (;;) {
                if (burst_ahead_ && inflight == 1 && pass_ + 2 * B < o.max_passes) {
                    ck(exc_burst_submit(ctx, B, dense ? 1 : 0), "burst ahead");
                    inflight = 2;
                }
                ck(exc_burst_collect(ctx, b_cnt.data(), b_l1.data(), b_linf.data(), nullptr), "burst");
                inflight--;
                const Faults f1 = faults();
                tot_maj_ += f1.maj - fb.maj;
                tot_min_ += f1.min - fb.min;
                converged = !report(f1.maj - fb.maj, f1.min - fb.min);
                fb = f1;
                if (converged || !inflight) break;
                // the burst ahead runs as this kind: read it and go back to the loop once the
                // demand picks the other kind or a measure has finished (the loop drops it)
                bool same = want_dense(p_cnt.data()) == dense;
                if (nm > 1 && o.drop_measures == "on")
                    for (int k = 0; k < nm; k++) same = same && p_linf[k] > M_.eps[k];
                if (!same) {
                    ck(exc_burst_collect(ctx, b_cnt.data(), b_l1.data(), b_linf.data(), nullptr), "burst ahead");
                    inflight--;
                    converged = !report(0, 0);
                    break;
                }
            }
            if (converged) {   // a burst still in flight starts from nothing pending: it changed nothing
                std::vector<uint64_t> scratch(b_cnt.size());
                while (inflight) { ck(exc_burst_collect(ctx, scratch.data(), nullptr, nullptr, nullptr), "burst ahead"); inflight--; }
            }
            maybe_enable_pull();
            if (converged) { stop_reason_ = dropped_global_ ? 1 : 0; break; }
            pass_--;   // the loop statement counts the last one
            continue;
        }

        int nres = 0, nstream = 0, ngather = 0;
        if (hot_skip && h_hotpend && o.hot_when != "end") launch_hot(0);
        uint64_t demand = 0;
        for (int c = 0; c < nc; c++) demand += h_cnt[2 * c + 1];
        if (pull_on_ && !hot_skip && ((double)demand >= pull_thr_ * (double)m || exc_dense_required(ctx) == 1)) {
            // dense pass over the whole cached graph (every chunk is resident once pull is on),
            // whatever the per-chunk list/range split would have been
            exc_chunk all{0, n, 0, m};
            ck(exc_push_resident(ctx, &all, 0, 0, 0), "dense pass");
            for (int c = 0; c < nc; c++) if (mode[c] == F_RESIDENT) { nres++; n_range_++; }
            n_launch_++;
            std::fill(mode.begin(), mode.end(), (int8_t)F_SKIP);
        }
        for (int c = 0; c < nc;) {   // resident chunks, runs merged into one launch
            if (mode[c] != F_RESIDENT) { c++; continue; }
            const bool ul = use_list(h_cnt[2 * c], chunks_[c].v1 - chunks_[c].v0, o.list_frac);
            uint64_t k = h_cnt[2 * c], e = h_cnt[2 * c + 1];
            int c1 = c + 1;
            if (push_kernel_ != EXC_KERNEL_GROUP)   // one launch for a run of chunks
                while (c1 < nc && mode[c1] == F_RESIDENT && use_list(h_cnt[2 * c1], chunks_[c1].v1 - chunks_[c1].v0, o.list_frac) == ul) {
                    k += h_cnt[2 * c1]; e += h_cnt[2 * c1 + 1]; c1++;
                }
            for (int j = c; j < c1; j++) { nres++; if (ul) n_list_++; else n_range_++; }
            exc_chunk run{chunks_[c].v0, chunks_[c1 - 1].v1, chunks_[c].e0, chunks_[c1 - 1].e1};
            ck(exc_push_resident(ctx, &run, ul, group_of(k, e), hot_skip), "resident push");
            n_launch_++;
            c = c1;
        }
        for (int c = 0; c < nc; c++) {
            if (mode[c] < F_STREAM) continue;
            const EdgeChunk& ch = chunks_[c];
            const int s = cur_;
            cur_ ^= 1;
            double w = 0;
            ck(exc_slot_acquire(ctx, s, &w), "slot acquire");
            t_wait_ += w;
            auto io0 = Clock::now();
            if (mode[c] == F_STREAM) {
                nstream++;
                g.read_edges(ch.e0, ch.e1, h_edges_[s]);
                if (!loaded_[c]) check_columns(h_edges_[s], ch.edges(), n, "a streamed chunk", threads_);
                bytes_stream_ += ch.edges() * sizeof(int32_t);
                t_io_ += secs(io0, Clock::now());
                const bool ul = use_list(h_cnt[2 * c], ch.v1 - ch.v0, o.list_frac);
                if (ul) n_list_++; else n_range_++;
                exc_chunk cc{ch.v0, ch.v1, ch.e0, ch.e1};
                ck(exc_stream_chunk(ctx, s, &cc, resident_[c] ? 1 : 0, 0, 0, ul, group_of(h_cnt[2 * c], h_cnt[2 * c + 1]), hot_skip), "stream push");
                n_launch_++;
                if (resident_[c]) loaded_[c] = 1;
            } else {
                // consecutive gathered chunks are staged together while their lists fit one
                // slot (the counts are exact: the snapshot's count and degree sum of a chunk
                // are its bitmap's pending vertices and their edges), so a pass with a few
                // pending vertices over many chunks makes one copy and one push, not one each
                int c1 = c + 1;
                uint64_t kk = h_cnt[2 * c], ee = h_cnt[2 * c + 1];
                while (c1 < nc && mode[c1] == F_GATHER && kk + h_cnt[2 * c1] <= (uint64_t)slot_verts_ &&
                       ee + h_cnt[2 * c1 + 1] <= slot_edges_) {
                    kk += h_cnt[2 * c1];
                    ee += h_cnt[2 * c1 + 1];
                    c1++;
                }
                ngather += c1 - c;
                const int k = gather_list(h_bitmap.data(), ch.v0, chunks_[c1 - 1].v1, h_vids_[s]);
                c = c1 - 1;
                const GatherStats gs = stage_lists(g, h_vids_[s], k, h_off_[s], h_edges_[s], o.dedup, threads_);
                const uint64_t ne = gs.edges;
                pass_pages += gs.pages;
                pass_runs += gs.runs;
                check_columns(h_edges_[s], ne, n, "a gathered chunk", threads_);
                // the reference's byte count of a gather: columns, ids and offsets (src/excent.cu:730)
                bytes_gather_ += ne * sizeof(int32_t) + (size_t)k * (sizeof(int32_t) + sizeof(uint64_t));
                t_io_ += secs(io0, Clock::now());
                const int G = o.group_opt ? o.group_opt : pick_group(k ? (double)ne / k : 1.0, gcap);
                ck(exc_gather_chunk(ctx, s, k, G), "gather push");
                if (k) n_launch_++;
            }
        }
        if (hot_skip && o.hot_when != "start") launch_hot(1);
        double w = 0;
        ck(exc_pass_sync(ctx, &w), "pass sync");
        t_wait_ += w;
        maybe_enable_pull();
        const Faults f1 = faults();
        tot_maj_ += f1.maj - f0.maj;
        tot_min_ += f1.min - f0.min;
        tot_pages_ += pass_pages;
        tot_runs_ += pass_runs;
        pass_line(pass_, po.pend_v, po.demand, nres, nstream, ngather, h_hotpend, f1.maj - f0.maj, f1.min - f0.min,
                  pass_pages, pass_runs, have_l1_, l1_, linf_);
    }
    T_iter_ = Clock::now();
    T_setup_ = T_begin;
    if (stop_reason_ == 2 && !batched()) {   // stopped by the pass limit: the residual of the final state
        ck(exc_snapshot(ctx, 0, h_cnt.data(), &h_hotpend), "final snapshot");
        have_l1_ = exc_residual_l1(ctx, l1_) == EXC_OK && exc_residual_max(ctx, linf_) == EXC_OK;
    }
}

// Reads an n-element device vector to the host in 16M-element slices, handing each slice
// (in vertex order) to `sink`.
template <class T, class Read, class Sink>
void for_each_slice(size_t n, Read read, Sink sink) {
    const size_t B = (size_t)1 << 24;
    std::vector<T> buf(std::min(B, std::max<size_t>(n, 1)));
    for (size_t v0 = 0; v0 < n; v0 += B) {
        const size_t cnt = std::min(B, n - v0);
        read(buf.data(), v0, cnt);
        sink(buf.data(), cnt);
    }
}

// Values (and --dump files, --trace-pending counters): src/excent.cu:761-789.  The sum of
// a measure is a double sum over its float values in vertex order; a batched run writes its
// K columns one after another to PREFIX.ppr.bin.
void Engine::results() {
    const Options& o = o_;
    const size_t n = (size_t)n_;
    const int ncols = batched() ? K_ : M_.nm;
    sum_.assign(ncols, 0.0);
    auto open_out = [&](const std::string& path) {
        FILE* f = fopen(path.c_str(), "wb");
        if (!f) die("cannot write " + path);
        return f;
    };
    FILE* shared = batched() && !o.dump.empty() ? open_out(o.dump + ".ppr.bin") : nullptr;
    for (int i = 0; i < ncols; i++) {
        FILE* f = shared;
        if (!batched() && !o.dump.empty()) f = open_out(o.dump + "." + MNAME[kinds_[i]] + ".bin");
        double& total = sum_[i];
        for_each_slice<float>(
            n, [&](float* dst, size_t v0, size_t cnt) { ck(exc_read_state(ctx_, i, 0, dst, v0, cnt), "result readback"); },
            [&](const float* xs, size_t cnt) {
                for (size_t j = 0; j < cnt; j++) total += xs[j];
                if (f && fwrite(xs, sizeof(float), cnt, f) != cnt) die("short write of the --dump output");
            });
        if (f && f != shared) fclose(f);
    }
    if (shared) fclose(shared);
    if (!o.trace_file.empty()) {
        FILE* f = open_out(o.trace_file);
        for_each_slice<uint16_t>(
            n, [&](uint16_t* dst, size_t v0, size_t cnt) { ck(exc_read_trace(ctx_, dst, v0, cnt), "trace readback"); },
            [&](const uint16_t* t, size_t cnt) {
                if (fwrite(t, sizeof(uint16_t), cnt, f) != cnt) die("short write of " + o.trace_file);
            });
        fclose(f);
    }
    T_end_ = Clock::now();
}

void Engine::report() {
    const Options& o = o_;
    FILE* out = out_;
    uint64_t pushed = 0;
    ck(exc_pushed(ctx_, &pushed), "pushed count");
    fprintf(out, "passes %d  measure %s (n=%d):", pass_, o.measure.c_str(), n_);
    if (batched()) {
        double tot = 0, lo = 1e300, hi = -1e300;
        for (double v : sum_) { tot += v; lo = std::min(lo, v); hi = std::max(hi, v); }
        fprintf(out, "  %d columns, ppr sum %.6e (column sums %.6e .. %.6e)", K_, tot, lo, hi);
    } else {
        for (int i = 0; i < M_.nm; i++) fprintf(out, "  %s sum %.6e", MNAME[kinds_[i]], sum_[i]);
    }
    fprintf(out, "\n");
    if (have_l1_) {   // extension: the FP64 global residual and the rule that ended the loop
        fprintf(out, "residual l1 (fp64):");
        for (int i = 0; i < M_.nm; i++) fprintf(out, "  %s %.6e (n*eps %.6e)", MNAME[kinds_[i]], l1_[i], (double)n_ * (double)M_.eps[i]);
        fprintf(out, " | stop: %s\n", stop_reason_ == 0 ? "nothing pending" : stop_reason_ == 1 ? "global l1 <= n*eps" : "pass limit");
        fprintf(out, "residual max:");
        for (int i = 0; i < M_.nm; i++) fprintf(out, "  %s %.6e (eps %.6e)", MNAME[kinds_[i]], (double)linf_[i], (double)M_.eps[i]);
        fprintf(out, "\n");
    }
    if (o.count_pushes) fprintf(out, "edges pushed: %.2f x |E|\n", m_ ? (double)pushed / m_ : 0.0);
    fprintf(out, "launches from GPU memory: %llu pending-only, %llu whole-range\n", (unsigned long long)n_list_, (unsigned long long)n_range_);
    fprintf(out, "bytes shipped: stream %.2f GB, gather %.2f GB (graph edges %.2f GB)\n", bytes_stream_ / 1e9, bytes_gather_ / 1e9, m_ * 4.0 / 1e9);
    fprintf(out, "page faults in passes: major %ld, minor %ld | gather dedup %s: %llu pages in %llu runs\n",
            tot_maj_, tot_min_, o.dedup ? "on" : "off", (unsigned long long)tot_pages_, (unsigned long long)tot_runs_);
    fprintf(out, "time: open %.2fs setup %.2fs passes %.2fs [cpu io %.2fs, gpu wait %.2fs, snapshot %.2fs] result %.2fs | internal total %.2fs\n",
            secs(T0_, T_open_), secs(T_open_, T_setup_), secs(T_setup_, T_iter_), t_io_, t_wait_, t_snap_,
            secs(T_iter_, T_end_), secs(T0_, T_end_));
    fflush(out);
}

void Engine::fill_stats(excent_run_stats* s) {
    memset(s, 0, sizeof *s);
    exc_stats st{};
    ck(exc_get_stats(ctx_, &st), "stats");
    s->time_passes_s = secs(T_setup_, T_iter_);
    s->time_snapshot_s = t_snap_;
    s->time_io_s = t_io_;
    s->time_wait_s = t_wait_;
    s->passes = pass_;
    s->pushed_edges = st.pushed_edges;
    s->bytes_stream = bytes_stream_;
    s->bytes_gather = bytes_gather_;
    s->launches = n_launch_;
    s->kernel = st;
    for (int i = 0; i < EXC_MAXM; i++) s->residual_l1[i] = have_l1_ && i < M_.nm ? l1_[i] : 0.0;
    s->stop_reason = stop_reason_;
    s->time_open_s = t_open_s_;
    s->time_setup_s = t_setup_s_;
    s->time_transpose_s = t_transpose_s_;
}

void Engine::write_stats_json() {
    if (o_.stats_json.empty()) return;
    FILE* j = fopen(o_.stats_json.c_str(), "w");
    if (!j) die("cannot write " + o_.stats_json);
    excent_run_stats rs;
    fill_stats(&rs);
    const exc_stats& st = rs.kernel;
    const double tp = rs.time_passes_s;
    fprintf(j, "{\"n\": %d, \"m\": %llu, \"chunks\": %d, \"passes\": %d, \"measure\": \"%s\", \"arch\": \"%s\", \"cus\": %d,\n",
            n_, (unsigned long long)m_, nc_, pass_, o_.measure.c_str(), arch_, cus_);
    fprintf(j, " \"time_open_s\": %.6f, \"time_setup_s\": %.6f, \"time_passes_s\": %.6f, \"time_io_s\": %.6f, \"time_wait_s\": %.6f, \"time_snapshot_s\": %.6f,\n",
            t_open_s_, t_setup_s_, tp, t_io_, t_wait_, t_snap_);
    fprintf(j, " \"time_transpose_s\": %.6f,\n", t_transpose_s_);
    fprintf(j, " \"pushed_edges\": %llu, \"pushed_vertices\": %llu, \"pushed_edge_measures\": %llu, \"gteps\": %.6f,\n",
            (unsigned long long)st.pushed_edges, (unsigned long long)st.pushed_vertices, (unsigned long long)st.pushed_edge_measures,
            tp > 0 ? st.pushed_edges / tp / 1e9 : 0.0);
    fprintf(j, " \"bytes_stream\": %llu, \"bytes_gather\": %llu, \"resident_chunks\": %d, \"hot_vertices\": %d, \"tau\": %.6f,\n",
            (unsigned long long)bytes_stream_, (unsigned long long)bytes_gather_, n_resident_, nh_, stream_ratio_);
    fprintf(j, " \"launches\": %llu, \"list_launches\": %llu, \"range_launches\": %llu,\n",
            (unsigned long long)n_launch_, (unsigned long long)n_list_, (unsigned long long)n_range_);
    fprintf(j, " \"kernel\": {\"snapshot_ms\": %.6f, \"snapshot_launches\": %llu, \"snapshot_bytes\": %.1f, \"push_ms\": %.6f, \"push_launches\": %llu, \"push_bytes\": %.1f, \"piece_ms\": %.6f, \"piece_launches\": %llu, \"piece_bytes\": %.1f, \"other_ms\": %.6f},\n",
            st.snapshot_ms, (unsigned long long)st.snapshot_launches, st.snapshot_bytes, st.push_ms, (unsigned long long)st.push_launches,
            st.push_bytes, st.piece_ms, (unsigned long long)st.piece_launches, st.piece_bytes, st.other_ms);
    if (have_l1_) {
        fprintf(j, " \"residual_l1\": {");
        for (int i = 0; i < M_.nm; i++) fprintf(j, "%s\"%s\": %.9e", i ? ", " : "", MNAME[kinds_[i]], l1_[i]);
        fprintf(j, "},\n");
    }
    fprintf(j, " \"stop\": \"%s\",\n", stop_reason_ == 0 ? "pending" : stop_reason_ == 1 ? "global" : "passes");
    fprintf(j, " \"sums\": {");
    if (batched()) {
        fprintf(j, "\"ppr\": [");
        for (size_t i = 0; i < sum_.size(); i++) fprintf(j, "%s%.9e", i ? ", " : "", sum_[i]);
        fprintf(j, "]");
    } else {
        for (int i = 0; i < M_.nm; i++) fprintf(j, "%s\"%s\": %.9e", i ? ", " : "", MNAME[kinds_[i]], sum_.size() > (size_t)i ? sum_[i] : 0.0);
    }
    fprintf(j, "}}\n");
    fclose(j);
}

}  // namespace excent

namespace {
thread_local std::string g_session_err;

template <class F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const excent::exit_error& e) {
        if (!e.msg.empty()) fprintf(stderr, "%s\n", e.msg.c_str());
        g_session_err = e.msg;
        fflush(stdout);
        return e.code ? e.code : 1;
    } catch (const std::exception& e) {
        fprintf(stderr, "%s\n", e.what());
        g_session_err = e.what();
        fflush(stdout);
        return 1;
    }
}
}  // namespace

extern "C" int excent_main(int argc, const char* const* argv) {
    int code = 0;
    int rc = guarded([&]() {
        excent::Options o;
        try {
            o = excent::parse(argc, argv);
        } catch (const excent::exit_error& e) {
            code = e.code;
            throw;
        }
        excent::Engine eng(o);
        eng.setup();
        eng.run_passes();
        eng.results();
        eng.report();
        eng.write_stats_json();
        return 0;
    });
    return rc ? rc : code;
}

extern "C" int excent_session_open(int argc, const char* const* argv, void** session) {
    *session = nullptr;
    return guarded([&]() {
        excent::Options o = excent::parse(argc, argv);
        std::unique_ptr<excent::Engine> eng(new excent::Engine(o, /*session=*/true));
        eng->setup();
        *session = eng.release();
        return 0;
    });
}

extern "C" int excent_session_run(void* session, int reinit, excent_run_stats* stats) {
    if (!session) return 1;
    return guarded([&]() {
        auto* eng = static_cast<excent::Engine*>(session);
        if (reinit) eng->reinit();
        eng->run_passes();
        if (stats) eng->fill_stats(stats);
        return 0;
    });
}

extern "C" int excent_session_set_column(void* session, int column) {
    if (!session) return 1;
    return guarded([&]() {
        static_cast<excent::Engine*>(session)->set_column(column);
        return 0;
    });
}

extern "C" int excent_session_values(void* session, int column, int which, float* out) {
    if (!session || !out) return 1;
    return guarded([&]() {
        static_cast<excent::Engine*>(session)->values(column, which, out);
        return 0;
    });
}

extern "C" int excent_session_vertices(void* session) {
    return session ? static_cast<excent::Engine*>(session)->vertices() : -1;
}

extern "C" int excent_session_results(void* session) {
    if (!session) return 1;
    return guarded([&]() {
        auto* eng = static_cast<excent::Engine*>(session);
        eng->results();
        eng->report();
        eng->write_stats_json();
        return 0;
    });
}

extern "C" int excent_session_close(void* session) {
    delete static_cast<excent::Engine*>(session);
    return 0;
}

extern "C" const char* excent_session_error(void) { return g_session_err.c_str(); }
