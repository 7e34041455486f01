// heat_solver_het.cpp -- command-line driver of the MI355X heat solver.
//
// Same argv, stdout banners and output files as the reference program
// src/heat_solver_het.cu:1340-1629 with one rank owning the whole mesh:
//   heat_solver_het N t_final dt field_prefix [outdir] [--flags]
// writing fi_trace_<tag>.csv, fi_summary_<tag>.csv, meta_<tag>.txt,
// T_final_<tag>_{clean,fault}_rank0.bin and, with --verify, verify.csv.
// All device work goes through the C ABI of include/heat.h.
//
// Added flags (no reference counterpart): --tol X and --max-iter N (the
// reference hard-codes 1e-8 and 500), --source-const Q, --ic sine,
// --precision fp64|fp32 (FP32-refined solve), --inner-tol X,
// --colored-impl gather|sweep|march|tile|tdma, --pcg-chunk K, --no-graph,
// --pcg-order fused|reference (summation order of the PCG inner products),
// --campaign FILE (many fault runs of one field in one process, below),
// --campaign-batch B (B of those runs advanced together as one batch).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "heat.h"

namespace {

struct ApiError {
    int code;
};

#define API(call)                                                                          \
    do {                                                                                   \
        int rc_ = (call);                                                                  \
        if (rc_ != HEAT_OK) {                                                              \
            fprintf(stderr, "[heat] %s failed (%d): %s\n", #call, rc_, heat_last_error()); \
            throw ApiError{rc_};                                                           \
        }                                                                                  \
    } while (0)

// one line of a --campaign file
struct CampaignRun {
    std::string dir;
    int level, step, bit;
    long target;
};

struct Opts {
    int N = 0;
    double t_final = 0, dt = 0;
    std::string field, outdir = "out", tag = "run";
    int fi_level = 0, fi_step = -1, fi_bit = 0, fi_rank = 0;
    long fi_target = 0;
    int snap_every = 5, verify = 0, kernel = HEAT_KERNEL_FAST, colored = 0, reproject = 0;
    int detect = 0, no_source = 0, ic_mode = 0, dump_snaps = 0, max_iter = 500;
    int source_mode = HEAT_SOURCE_GAUSSIAN;
    double source_value = 0.0, tol = 1e-8, thr_user = 1e-13;
    std::string diag_cg;
    int mixed = 0;
    double inner_tol = 1e-4;
    int colored_impl = -1, pcg_chunk = 8, use_graph = 1, pcg_order = HEAT_PCG_FUSED;
    std::string campaign;
    std::vector<CampaignRun> campaign_runs;
    int campaign_batch = 0;
};

// Lines "DIR LEVEL STEP TARGET BIT"; DIR is everything before the last four
// fields, so it may hold spaces. False (with a message) on a bad file.
bool read_campaign(const std::string& path, std::vector<CampaignRun>& runs) {
    FILE* f = fopen(path.c_str(), "r");
    if (!f) {
        fprintf(stderr, "cannot read %s\n", path.c_str());
        return false;
    }
    char line[4096];
    bool bad_line = false;
    while (fgets(line, sizeof line, f)) {
        std::string t(line);
        while (!t.empty() && (t.back() == '\n' || t.back() == '\r' || t.back() == ' ')) t.pop_back();
        if (t.empty()) continue;
        size_t cut = t.size();
        for (int k = 0; k < 4 && cut != std::string::npos; ++k) {
            cut = t.find_last_of(' ', cut - 1);
            if (cut == 0) cut = std::string::npos;
        }
        CampaignRun r;
        char tail;
        if (cut == std::string::npos ||
            sscanf(t.c_str() + cut, "%d %d %ld %d %c", &r.level, &r.step, &r.target, &r.bit, &tail) != 4) {
            bad_line = true;
            break;
        }
        r.dir = t.substr(0, cut);
        runs.push_back(r);
    }
    fclose(f);
    if (bad_line) fprintf(stderr, "%s: every line must read DIR LEVEL STEP TARGET BIT\n", path.c_str());
    return !bad_line;
}

void usage(const char* prog) {
    fprintf(stderr,
            "usage: %s N t_final dt field_prefix [outdir]\n"
            "   [--fi LEVEL STEP TARGET BIT] [--fi-rank R] [--snap S] [--tag NAME]\n"
            "   [--kernel gauss|fast] [--scatter atomic|colored] [--thr X]\n"
            "   [--reproject-bc] [--detect] [--verify] [--no-source] [--ic zero|gaussian|sine] [--dump-snaps]\n"
            "   [--diag-cg FILE]   (one CG solve, logs iteration,rel_residual to FILE, then exits)\n"
            "   [--tol X] [--max-iter N] [--source-const Q] [--precision fp64|fp32] [--inner-tol X]\n"
            "   [--colored-impl gather|sweep|march|tile|tdma] [--pcg-chunk K] [--no-graph]\n"
            "   [--pcg-order fused|reference] [--campaign FILE] [--campaign-batch B]\n"
            "  LEVEL: 1=G global memory, 2=R accumulator, 3=H halo staging buffer\n"
            "  BIT:   0-51 mantissa, 52-62 exponent, 63 sign\n",
            prog);
}

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct RunResult {
    std::vector<int> snap_steps, cg_iters;
    std::vector<double> E, spread;
    double solve_ms = 0.0;
    int nonfinite = 0;
    std::vector<double> d1_abft, d2_resid, d3_tmin, d3_tmax;
    double d1_max = 0, d2_max = 0;
    int d1_first = -1, d2_first = -1, d3_first = -1;
};

struct Solver {
    heat_op* op = nullptr;
    Opts o;
    long nn = 0;
    double stats[4] = {0, 0, 0, 0};
    double *T = nullptr, *b = nullptr, *r = nullptr, *Ap = nullptr;
    double chk_tol = 1e-10;

    int pcg(double* x, const double* rhs) {
        int it = 0;
        if (o.mixed) {
            int outer = 0, inner = 0;
            API(heat_pcg_solve_mixed(op, rhs, x, 50, o.max_iter, o.tol, o.inner_tol, &outer, &inner));
            it = inner;
        } else {
            API(heat_pcg_solve(op, rhs, x, o.max_iter, o.tol, &it, nullptr));
        }
        return it;
    }
    double norm2(const double* v) {
        double s = 0;
        API(heat_norm2(op, v, &s));
        return s;
    }
    void snapshot(std::vector<double>& out) {
        out.resize(nn);
        API(heat_download(op, out.data(), T, nn));
    }
};

// global_rel_err, src/heat_solver_het.cu:1054-1080 (one rank: every node owned)
double global_rel_err(const std::vector<double>& a, const std::vector<double>& b, double* spread) {
    double num = 0.0, den = 0.0, dmax = 0.0;
    const long n = (long)a.size();
    for (long i = 0; i < n; ++i) {
        const double d = a[i] - b[i];
        num += d * d;
        den += b[i] * b[i];
        if (std::fabs(d) > dmax) dmax = std::fabs(d);
    }
    if (spread) {
        long cnt = 0;
        for (long i = 0; i < n; ++i)
            if (std::fabs(a[i] - b[i]) > 1e-3 * dmax) ++cnt;
        *spread = (dmax > 0 && n > 0) ? (double)cnt / (double)n : 0.0;
    }
    if (den < 1e-300) den = 1.0;
    return std::sqrt(num / den);
}

// run_case, src/heat_solver_het.cu:1107-1259
void run_case(Solver& S, int n_steps, int level, int step_fi, long target, int bit,
              std::vector<std::vector<double>>* keep, const std::vector<std::vector<double>>* ref,
              RunResult& res, const char* case_tag, const RunResult* d3_ref) {
    const Opts& o = S.o;
    res = RunResult();
    API(heat_initial_condition(S.op, S.T, o.ic_mode, 0.0, S.stats[0] / S.stats[2]));
    API(heat_disarm_fault(S.op));
    if (keep) keep->clear();
    std::vector<double> tmp;
    size_t si = 0;
    for (int step = 0; step < n_steps; ++step) {
        if (level > 0 && step == step_fi) {
            if (level == 1) {
                if (o.fi_rank == 0) API(heat_flip_bit(S.op, S.T, target % S.nn, bit));
            } else {
                API(heat_arm_fault(S.op, level, target, bit));
            }
        }
        API(heat_synchronize(S.op));
        const double t0 = now_ms();
        API(heat_build_rhs(S.op, S.T, S.b));
        const int it = S.pcg(S.T, S.b);
        if (o.reproject) API(heat_set_dirichlet(S.op, S.T, 0.0));
        API(heat_synchronize(S.op));
        res.solve_ms += now_ms() - t0;
        API(heat_disarm_fault(S.op));
        res.cg_iters.push_back(it);
        if (o.detect) {
            double d1 = 0;
            API(heat_abft_take(S.op, &d1, nullptr));                 // abft_step_max of the step
            API(heat_apply_operator(S.op, S.T, S.Ap, o.dt, 0, 1));   // D2: fresh clean matvec
            API(heat_copy(S.op, S.b, S.r));
            API(heat_axpy(S.op, -1.0, S.Ap, S.r));
            double bn = S.norm2(S.b);
            if (bn < 1e-300) bn = 1.0;
            const double d2 = S.norm2(S.r) / bn;
            double tmin = 0, tmax = 0;
            API(heat_minmax(S.op, S.T, &tmin, &tmax));
            bool d3_fire = !std::isfinite(tmin) || !std::isfinite(tmax);
            if (d3_ref && step < (int)d3_ref->d3_tmax.size()) {
                const double ref_max = d3_ref->d3_tmax[step], ref_min = d3_ref->d3_tmin[step];
                const double margin = 1e-6 * std::fmax(std::fabs(ref_max), 1e-300) + 1e-12;
                if (tmax > ref_max + margin || tmin < ref_min - margin) d3_fire = true;
            }
            res.d1_abft.push_back(d1);
            res.d2_resid.push_back(d2);
            res.d3_tmin.push_back(tmin);
            res.d3_tmax.push_back(tmax);
            if (!(d1 <= S.chk_tol) && res.d1_first < 0) res.d1_first = step;
            if (!(d2 <= 1e-6) && res.d2_first < 0) res.d2_first = step;
            if (d3_fire && res.d3_first < 0) res.d3_first = step;
            res.d1_max = std::isfinite(d1) ? std::fmax(res.d1_max, d1) : 1e300;
            res.d2_max = std::isfinite(d2) ? std::fmax(res.d2_max, d2) : 1e300;
            API(heat_abft_take(S.op, nullptr, nullptr));             // abft_step_max = 0
        }
        const bool snap = (o.snap_every > 0 && step % o.snap_every == 0) || step == n_steps - 1;
        if (snap) {
            if (keep) {
                keep->emplace_back();
                S.snapshot(keep->back());
            }
            if (o.dump_snaps) {
                if (!keep) S.snapshot(tmp);
                const std::vector<double>& v = keep ? keep->back() : tmp;
                char pth[1024];
                snprintf(pth, sizeof pth, "%s/snap_%s_%s_%05d_rank%d.bin", o.outdir.c_str(), o.tag.c_str(),
                         case_tag, step, 0);
                FILE* f = fopen(pth, "wb");
                if (f) {
                    fwrite(v.data(), sizeof(double), v.size(), f);
                    fclose(f);
                }
            }
            if (ref) {
                S.snapshot(tmp);
                double sp = 0;
                const double E = global_rel_err(tmp, (*ref)[si], &sp);
                if (!std::isfinite(E)) res.nonfinite = 1;
                res.E.push_back(E);
                res.spread.push_back(sp);
            }
            res.snap_steps.push_back(step);
            ++si;
        }
    }
}

bool write_field(const std::vector<double>& T, const std::string& outdir, const std::string& tag) {
    char path[1024];
    snprintf(path, sizeof path, "%s/T_final_%s_rank0.bin", outdir.c_str(), tag.c_str());
    FILE* f = fopen(path, "wb");
    if (!f) {
        fprintf(stderr, "cannot write %s\n", path);
        return false;
    }
    fwrite(T.data(), sizeof(double), T.size(), f);
    fclose(f);
    return true;
}

// Field statistics and labels shared by every case of one invocation.
struct CaseLabels {
    int n_steps, nprocs;
    const char *kname, *sname;
    double kmin, kmax, rmin, rmax;
};

// fi_trace_<tag>.csv, fi_summary_<tag>.csv and meta_<tag>.txt of one twin run
// into DIR, and its stdout summary (src/heat_solver_het.cu:1520-1621). o holds
// this run's fault (fi_level 0: a clean run without F). False when a file
// cannot be written.
bool write_case(const Opts& o, const std::string& dir, const CaseLabels& c, const RunResult& rA,
                const RunResult& rB, const RunResult& rF, int on_bnd) {
    const bool fault = o.fi_level > 0;
    double Efl_max = 0, Efa_max = 0, Efa_int = 0, Efa_final = 0, Efl_final = 0, spread_at_max = 0,
           spread_final = 0;
    int step_max = -1, iters_delta_max = 0;
    for (size_t s = 0; s < rB.E.size(); ++s) {
        Efl_max = std::fmax(Efl_max, rB.E[s]);
        Efl_final = rB.E[s];
        if (fault) {
            const double E = rF.E[s];
            if (std::isfinite(E) && E > Efa_max) {
                Efa_max = E;
                step_max = rB.snap_steps[s];
                spread_at_max = rF.spread[s];
            }
            Efa_final = E;
            spread_final = rF.spread[s];
            const double w = (s + 1 < rB.E.size()) ? (rB.snap_steps[s + 1] - rB.snap_steps[s]) * o.dt : 0.0;
            if (std::isfinite(E)) Efa_int += E * w;
        }
    }
    for (size_t s = 0; fault && s < rF.cg_iters.size(); ++s)
        iters_delta_max = std::max(iters_delta_max, rF.cg_iters[s] - rA.cg_iters[s]);
    const double thr = std::fmax(o.thr_user, 10.0 * Efl_max);
    const char* cls = "clean";
    if (fault) {
        if (rF.nonfinite) cls = "detected";
        else if (Efa_max <= thr) cls = "benign";
        else if (Efa_final <= thr) cls = "transient";
        else cls = "persistent";
    }
    const RunResult& rX = fault ? rF : rA;
    const std::string tp = dir + "/fi_trace_" + o.tag + ".csv";
    FILE* tf = fopen(tp.c_str(), "w");
    if (!tf) {
        fprintf(stderr, "cannot write %s\n", tp.c_str());
        return false;
    }
    fprintf(tf, "step,t,E_noise_floor,E_fault,spread_F,cg_A,cg_B,cg_F,d1_abft_A,d1_abft_F,d2_resid_A,"
                "d2_resid_F,d3_tmin_F,d3_tmax_F,hyb_block_err_A,hyb_block_err_F,hyb_d4_F\n");
    for (size_t s = 0; s < rB.E.size(); ++s) {
        const int st = rB.snap_steps[s];
        fprintf(tf, "%d,%.6e,%.10e,%.10e,%.4f,%d,%d,%d", st, st * o.dt, rB.E[s], fault ? rF.E[s] : 0.0,
                fault ? rF.spread[s] : 0.0, rA.cg_iters[st], rB.cg_iters[st], fault ? rF.cg_iters[st] : -1);
        if (o.detect)
            fprintf(tf, ",%.3e,%.3e,%.3e,%.3e,%.6e,%.6e", rA.d1_abft[st], rX.d1_abft[st], rA.d2_resid[st],
                    rX.d2_resid[st], rX.d3_tmin[st], rX.d3_tmax[st]);
        else
            fprintf(tf, ",,,,,,");
        fprintf(tf, ",,,\n");
    }
    fclose(tf);
    double meanA = 0;
    for (int it : rA.cg_iters) meanA += it;
    meanA /= (double)std::max<size_t>(1, rA.cg_iters.size());
    const std::string sp = dir + "/fi_summary_" + o.tag + ".csv";
    FILE* sf = fopen(sp.c_str(), "w");
    if (!sf) {
        fprintf(stderr, "cannot write %s\n", sp.c_str());
        return false;
    }
    fprintf(sf, "tag,N,dt,n_steps,nprocs,kernel,scatter,reproject_bc,kmin,kmax,rhocmin,rhocmax,"
                "fi_level,fi_rank,fi_step,fi_target,fi_bit,target_on_boundary,"
                "E_floor_max,E_floor_final,E_fault_max,step_at_max,E_fault_int,E_fault_final,"
                "thr,iters_delta_max,mean_cg_iters_A,solve_ms_A,solve_ms_B,solve_ms_F,class,spread_at_max,spread_final,"
                "detect,d1_first,d2_first,d3_first,d1_max_F,d2_max_F,d1_max_A,d2_max_A,detected_by,"
                "hybrid,hyb_block_err_max_A,hyb_block_err_max_F,d4_first_F\n");
    fprintf(sf, "%s,%d,%.3e,%d,%d,%s,%s,%d,%g,%g,%g,%g,%d,%d,%d,%ld,%d,%d,"
                "%.6e,%.6e,%.6e,%d,%.6e,%.6e,%.3e,%d,%.2f,%.3f,%.3f,%.3f,%s",
            o.tag.c_str(), o.N, o.dt, c.n_steps, c.nprocs, c.kname, c.sname, o.reproject, c.kmin, c.kmax, c.rmin, c.rmax,
            o.fi_level, o.fi_rank, o.fi_step, o.fi_target, o.fi_bit, on_bnd, Efl_max, Efl_final, Efa_max,
            step_max, Efa_int, Efa_final, thr, iters_delta_max, meanA, rA.solve_ms, rB.solve_ms,
            rF.solve_ms, cls);
    fprintf(sf, ",%.4f,%.4f", spread_at_max, spread_final);
    std::string by;
    if (o.detect) {
        if (rX.d1_first >= 0) by += "D1";
        if (rX.d2_first >= 0) by += (by.empty() ? "" : "+") + std::string("D2");
        if (rX.d3_first >= 0) by += (by.empty() ? "" : "+") + std::string("D3");
        if (by.empty()) by = "none";
    } else {
        by = "off";
    }
    fprintf(sf, ",%d,%d,%d,%d,%.3e,%.3e,%.3e,%.3e,%s", o.detect, rX.d1_first, rX.d2_first, rX.d3_first,
            rX.d1_max, rX.d2_max, rA.d1_max, rA.d2_max, by.c_str());
    fprintf(sf, ",%d,%.3e,%.3e,%d\n", 0, 0.0, 0.0, -1);
    if (o.detect)
        printf("detectors    D1 first %d (max %.2e)  D2 first %d (max %.2e)  D3 first %d  -> %s   "
               "[clean A: D1 %.2e D2 %.2e]\n",
               rX.d1_first, rX.d1_max, rX.d2_first, rX.d2_max, rX.d3_first, by.c_str(), rA.d1_max, rA.d2_max);
    fclose(sf);
    printf("\n=== Done (%s) ===\n", o.tag.c_str());
    printf("solve wall A/B/F (ms): %.1f / %.1f / %.1f   mean CG iters A: %.2f\n", rA.solve_ms, rB.solve_ms,
           rF.solve_ms, meanA);
    printf("noise floor  max %.3e  final %.3e\n", Efl_max, Efl_final);
    if (fault)
        printf("fault        max %.3e @step %d  int %.3e  final %.3e  iters+%d  bnd %d  -> %s\n", Efa_max,
               step_max, Efa_int, Efa_final, iters_delta_max, on_bnd, cls);
    const std::string mp = dir + "/meta_" + o.tag + ".txt";
    FILE* mf = fopen(mp.c_str(), "w");
    if (mf) {
        fprintf(mf, "N %d\nnprocs %d\ndt %g\nn_steps %d\nfield %s\nkernel %s\nscatter %s\n", o.N, c.nprocs,
                o.dt, c.n_steps, o.field.c_str(), c.kname, c.sname);
        fclose(mf);
    }
    return true;
}

bool is_boundary(int N, int ix, int iy, int iz) {
    return ix == 0 || ix == N || iy == 0 || iy == N || iz == 0 || iz == N;
}

// target_on_boundary, src/heat_solver_het.cu:1272-1290
int target_on_boundary(const Opts& o) {
    const int N = o.N, nn1 = N + 1;
    if (o.fi_level == 1) {
        const long nd = o.fi_target % ((long)nn1 * nn1 * nn1);
        return is_boundary(N, (int)(nd / ((long)nn1 * nn1)), (int)((nd / nn1) % nn1), (int)(nd % nn1));
    }
    if (o.fi_level == 2) {
        const long e = o.fi_target % ((long)N * N * N);
        return is_boundary(N, (int)(e / ((long)N * N)), (int)((e / N) % N), (int)(e % N));
    }
    return 0;
}

bool load_fields(const Opts& o, std::vector<double>& k, std::vector<double>& rhoc) {
    const long ne = (long)o.N * o.N * o.N;
    const std::string fk = o.field + "_k.bin", fr = o.field + "_rhoc.bin";
    FILE* f1 = fopen(fk.c_str(), "rb");
    FILE* f2 = fopen(fr.c_str(), "rb");
    if (!f1 || !f2) {
        fprintf(stderr, "Cannot open %s / %s (run gen_field.py)\n", fk.c_str(), fr.c_str());
        if (f1) fclose(f1);
        if (f2) fclose(f2);
        return false;
    }
    k.resize(ne);
    rhoc.resize(ne);
    const bool bad = fread(k.data(), sizeof(double), ne, f1) != (size_t)ne ||
                     fread(rhoc.data(), sizeof(double), ne, f2) != (size_t)ne;
    fclose(f1);
    fclose(f2);
    if (bad) {
        fprintf(stderr, "Field file too small: expected %ld doubles for this rank "
                        "(grid mismatch? regenerate with gen_field.py --n %d)\n", ne, o.N);
        return false;
    }
    return true;
}

int parse(int argc, char** argv, Opts& o) {
    if (argc < 5) {
        usage(argv[0]);
        return 1;
    }
    o.N = atoi(argv[1]);
    o.t_final = atof(argv[2]);
    o.dt = atof(argv[3]);
    o.field = argv[4];
    for (int i = 5; i < argc; ++i) {
        const char* a = argv[i];
        auto has = [&](int k) { return i + k < argc; };
        if (!strcmp(a, "--fi") && has(4)) {
            o.fi_level = atoi(argv[++i]);
            o.fi_step = atoi(argv[++i]);
            o.fi_target = atol(argv[++i]);
            o.fi_bit = atoi(argv[++i]);
        } else if (!strcmp(a, "--fi-rank") && has(1)) { o.fi_rank = atoi(argv[++i]);
        } else if (!strcmp(a, "--snap") && has(1)) { o.snap_every = atoi(argv[++i]);
        } else if (!strcmp(a, "--tag") && has(1)) { o.tag = argv[++i];
        } else if (!strcmp(a, "--thr") && has(1)) { o.thr_user = atof(argv[++i]);
        } else if (!strcmp(a, "--kernel") && has(1)) {
            o.kernel = strcmp(argv[++i], "gauss") != 0 ? HEAT_KERNEL_FAST : HEAT_KERNEL_GAUSS;
        } else if (!strcmp(a, "--scatter") && has(1)) { o.colored = !strcmp(argv[++i], "colored");
        } else if (!strcmp(a, "--reproject-bc")) { o.reproject = 1;
        } else if (!strcmp(a, "--detect")) { o.detect = 1;
        } else if (!strcmp(a, "--diag-cg") && has(1)) { o.diag_cg = argv[++i];
        } else if (!strcmp(a, "--no-source")) { o.no_source = 1;
        } else if (!strcmp(a, "--dump-snaps")) { o.dump_snaps = 1;
        } else if (!strcmp(a, "--ic") && has(1)) {
            ++i;
            o.ic_mode = !strcmp(argv[i], "gaussian") ? 1 : !strcmp(argv[i], "sine") ? 2 : 0;
        } else if (!strcmp(a, "--verify")) { o.verify = 1;
        } else if (!strcmp(a, "--tol") && has(1)) { o.tol = atof(argv[++i]);
        } else if (!strcmp(a, "--max-iter") && has(1)) { o.max_iter = atoi(argv[++i]);
        } else if (!strcmp(a, "--source-const") && has(1)) {
            o.source_mode = HEAT_SOURCE_CONST;
            o.source_value = atof(argv[++i]);
        } else if (!strcmp(a, "--precision") && has(1)) { o.mixed = !strcmp(argv[++i], "fp32");
        } else if (!strcmp(a, "--inner-tol") && has(1)) { o.inner_tol = atof(argv[++i]);
        } else if (!strcmp(a, "--colored-impl") && has(1)) {
            const char* v = argv[++i];
            o.colored_impl = !strcmp(v, "sweep") ? HEAT_COLORED_SWEEP : !strcmp(v, "march") ? HEAT_COLORED_MARCH
                           : !strcmp(v, "tile") ? HEAT_COLORED_TILE : !strcmp(v, "tdma") ? HEAT_COLORED_TDMA
                           : HEAT_COLORED_GATHER;
        } else if (!strcmp(a, "--pcg-order") && has(1)) {
            o.pcg_order = !strcmp(argv[++i], "reference") ? HEAT_PCG_REFERENCE_ORDER : HEAT_PCG_FUSED;
        } else if (!strcmp(a, "--pcg-chunk") && has(1)) { o.pcg_chunk = atoi(argv[++i]);
        } else if (!strcmp(a, "--no-graph")) { o.use_graph = 0;
        } else if (!strcmp(a, "--campaign") && has(1)) { o.campaign = argv[++i];
        } else if (!strcmp(a, "--campaign-batch") && has(1)) { o.campaign_batch = atoi(argv[++i]);
        } else if (!strcmp(a, "--qrank") || !strcmp(a, "--qblock")) {
            fprintf(stderr, "%s: the hybrid quantum-rank mode is not part of this solver\n", a);
            return 1;
        } else if (a[0] != '-') { o.outdir = a;
        } else {
            fprintf(stderr, "unknown option %s\n", a);
            usage(argv[0]);
            return 1;
        }
    }
    const int nprocs = 1;
    if (o.fi_level > 0 && (o.fi_rank < 0 || o.fi_rank >= nprocs)) {
        fprintf(stderr, "--fi-rank must be in [0,%d)\n", nprocs);
        return 1;
    }
    if (o.fi_level == 3) {
        fprintf(stderr, "level H needs >= 2 ranks and --fi-rank < nprocs-1\n");
        return 1;
    }
    if (o.fi_level == 4) {
        fprintf(stderr, "level 4 (handoff transit) needs --qrank --qblock\n");
        return 1;
    }
    if (o.fi_level > 0 && (o.fi_bit < 0 || o.fi_bit > 63)) {
        fprintf(stderr, "BIT must be in 0..63\n");
        return 1;
    }
    if (!o.campaign.empty() && (o.fi_level > 0 || o.dump_snaps || o.verify || !o.diag_cg.empty())) {
        fprintf(stderr, "--campaign takes its faults from FILE and runs without --fi, --dump-snaps, --verify "
                        "and --diag-cg\n");
        return 1;
    }
    if (!o.campaign.empty() && !read_campaign(o.campaign, o.campaign_runs)) return 1;
    return 0;
}

// --campaign FILE: the fault runs of one field in one process (the lines of
// FILE, read by read_campaign). The clean runs A and B are the same
// for every fault of the field (the solve is deterministic), so they run once;
// each fault run F then writes the files of a single invocation with
// --fi LEVEL STEP TARGET BIT into DIR (which must exist), except the T_final
// dumps. This is the process-per-run loop of scripts/ensemble/run_ensemble.py:78-104
// with one process per field. A line that a single invocation would reject
// (level other than 1 or 2, BIT outside 0..63) is reported and skipped, with
// no files written, and the return status is 1.
int run_campaign(Solver& S, int n_steps, const CaseLabels& lab, const std::vector<std::vector<double>>& snapA,
                 const RunResult& rA, const RunResult& rB) {
    const std::vector<CampaignRun>& runs = S.o.campaign_runs;
    int status = 0;
    const double t0 = now_ms();
    for (size_t i = 0; i < runs.size(); ++i) {
        const CampaignRun& r = runs[i];
        if ((r.level != 1 && r.level != 2) || r.bit < 0 || r.bit > 63) {
            fprintf(stderr, "[campaign] %s: level %d bit %d cannot run on one rank\n", r.dir.c_str(), r.level, r.bit);
            status = 1;
            continue;
        }
        Opts oc = S.o;
        oc.campaign_runs.clear();
        oc.fi_level = r.level;
        oc.fi_step = r.step;
        oc.fi_target = r.target;
        oc.fi_bit = r.bit;
        RunResult rF;
        run_case(S, n_steps, r.level, r.step, r.target, r.bit, nullptr, &snapA, rF, "fault", &rA);
        if (!write_case(oc, r.dir, lab, rA, rB, rF, target_on_boundary(oc))) status = 1;
        printf("[campaign] %zu/%zu %s  %.1f s\n", i + 1, runs.size(), r.dir.c_str(), (now_ms() - t0) / 1e3);
        fflush(stdout);
    }
    return status;
}

// One member of a batched campaign: a batch handle and its temperature.
struct Member {
    heat_op* op;
    double* T;
};

// run_case for the fault runs RUNS[0..m) on the first m members of BATCH:
// the same per-step sequence (the flip or the arming, b = M T + dt F and the
// PCG solve, the detectors, the snapshots against the clean run A), with
// the solves of all members advanced by one heat_batch_step.  A member's
// solve_ms is the batch step's wall time divided among the members.
void run_batch_cases(Solver& S, heat_batch* batch, std::vector<Member>& mem, int m, const CampaignRun* const* runs,
                     int n_steps, const std::vector<std::vector<double>>& snapA, const RunResult& rA,
                     std::vector<RunResult>& res) {
    const Opts& o = S.o;
    const int B = (int)mem.size();
    std::vector<int> active(B, 0);
    for (int i = 0; i < m; ++i) active[i] = 1;
    API(heat_batch_set_active(batch, active.data()));
    res.assign(m, RunResult());
    for (int i = 0; i < m; ++i) {
        API(heat_initial_condition(mem[i].op, mem[i].T, o.ic_mode, 0.0, S.stats[0] / S.stats[2]));
        API(heat_disarm_fault(mem[i].op));
    }
    std::vector<int> iters(B, 0);
    std::vector<double> det_d1(B), det_d2(B), det_tmin(B), det_tmax(B);
    std::vector<double> tmp;
    size_t si = 0;
    for (int step = 0; step < n_steps; ++step) {
        for (int i = 0; i < m; ++i) {
            const CampaignRun& r = *runs[i];
            if (step != r.step) continue;
            if (r.level == 1) {
                if (o.fi_rank == 0) API(heat_flip_bit(mem[i].op, mem[i].T, r.target % S.nn, r.bit));
            } else {
                API(heat_arm_fault(mem[i].op, r.level, r.target, r.bit));
            }
        }
        for (int i = 0; i < m; ++i) API(heat_synchronize(mem[i].op));
        const double t0 = now_ms();
        API(heat_batch_step(batch, 1, o.max_iter, o.tol, iters.data()));
        for (int i = 0; i < m && o.reproject; ++i) API(heat_set_dirichlet(mem[i].op, mem[i].T, 0.0));
        for (int i = 0; i < m; ++i) API(heat_synchronize(mem[i].op));
        const double share = (now_ms() - t0) / m;
        const bool snap = (o.snap_every > 0 && step % o.snap_every == 0) || step == n_steps - 1;
        // the detectors of all members at once: the values of run_case's
        // sequence (heat_batch_detect)
        if (o.detect) API(heat_batch_detect(batch, det_d1.data(), det_d2.data(), det_tmin.data(), det_tmax.data()));
        for (int i = 0; i < m; ++i) {
            RunResult& rr = res[i];
            heat_op* op = mem[i].op;
            rr.solve_ms += share;
            API(heat_disarm_fault(op));
            rr.cg_iters.push_back(iters[i]);
            if (o.detect) {
                const double d1 = det_d1[i], d2 = det_d2[i], tmin = det_tmin[i], tmax = det_tmax[i];
                bool d3_fire = !std::isfinite(tmin) || !std::isfinite(tmax);
                if (step < (int)rA.d3_tmax.size()) {
                    const double ref_max = rA.d3_tmax[step], ref_min = rA.d3_tmin[step];
                    const double margin = 1e-6 * std::fmax(std::fabs(ref_max), 1e-300) + 1e-12;
                    if (tmax > ref_max + margin || tmin < ref_min - margin) d3_fire = true;
                }
                rr.d1_abft.push_back(d1);
                rr.d2_resid.push_back(d2);
                rr.d3_tmin.push_back(tmin);
                rr.d3_tmax.push_back(tmax);
                if (!(d1 <= S.chk_tol) && rr.d1_first < 0) rr.d1_first = step;
                if (!(d2 <= 1e-6) && rr.d2_first < 0) rr.d2_first = step;
                if (d3_fire && rr.d3_first < 0) rr.d3_first = step;
                rr.d1_max = std::isfinite(d1) ? std::fmax(rr.d1_max, d1) : 1e300;
                rr.d2_max = std::isfinite(d2) ? std::fmax(rr.d2_max, d2) : 1e300;
            }
            if (snap) {
                tmp.resize(S.nn);
                API(heat_download(op, tmp.data(), mem[i].T, S.nn));
                double sp = 0;
                const double E = global_rel_err(tmp, snapA[si], &sp);
                if (!std::isfinite(E)) rr.nonfinite = 1;
                rr.E.push_back(E);
                rr.spread.push_back(sp);
                rr.snap_steps.push_back(step);
            }
        }
        if (snap) ++si;
    }
}

// --campaign with --campaign-batch B: the fault runs in groups of B members
// of one batch of handles (same field and configuration as the clean runs).
int run_campaign_batched(Solver& S, const heat_config& cfg, const std::vector<double>& hk,
                         const std::vector<double>& hr, int n_steps, const CaseLabels& lab,
                         const std::vector<std::vector<double>>& snapA, const RunResult& rA, const RunResult& rB) {
    const std::vector<CampaignRun>& runs = S.o.campaign_runs;
    int status = 0;
    std::vector<const CampaignRun*> todo;
    for (const CampaignRun& r : runs) {
        if ((r.level != 1 && r.level != 2) || r.bit < 0 || r.bit > 63) {
            fprintf(stderr, "[campaign] %s: level %d bit %d cannot run on one rank\n", r.dir.c_str(), r.level, r.bit);
            status = 1;
            continue;
        }
        todo.push_back(&r);
    }
    const int B = std::min<int>(S.o.campaign_batch, (int)todo.size());
    if (B <= 0) return status;
    std::vector<heat_config> cfgs(B, cfg);
    std::vector<const double*> kp(B, hk.data()), rp(B, hr.data());
    heat_batch* batch = nullptr;
    API(heat_batch_setup(B, cfgs.data(), kp.data(), rp.data(), &batch));
    API(heat_batch_set_tuning(batch, S.o.pcg_chunk, S.o.use_graph));
    std::vector<Member> mem(B);
    for (int i = 0; i < B; ++i) {
        API(heat_batch_instance(batch, i, &mem[i].op));
        API(heat_vector(mem[i].op, HEAT_VEC_T, &mem[i].T));
    }
    const double t0 = now_ms();
    for (size_t first = 0; first < todo.size(); first += B) {
        const int m = (int)std::min<size_t>(B, todo.size() - first);
        std::vector<RunResult> res;
        run_batch_cases(S, batch, mem, m, todo.data() + first, n_steps, snapA, rA, res);
        for (int i = 0; i < m; ++i) {
            const CampaignRun& r = *todo[first + i];
            Opts oc = S.o;
            oc.campaign_runs.clear();
            oc.fi_level = r.level;
            oc.fi_step = r.step;
            oc.fi_target = r.target;
            oc.fi_bit = r.bit;
            if (!write_case(oc, r.dir, lab, rA, rB, res[i], target_on_boundary(oc))) status = 1;
        }
        printf("[campaign] %zu/%zu runs (batches of %d)  %.1f s\n", first + m, todo.size(), B,
               (now_ms() - t0) / 1e3);
        fflush(stdout);
    }
    API(heat_batch_cleanup(batch));
    return status;
}

int run(const Opts& o) {
    const int nprocs = 1;
    const int n_steps = (int)(o.t_final / o.dt + 0.5);
    const char* kname = o.kernel == HEAT_KERNEL_GAUSS ? "gauss" : "fast";
    const char* sname = o.colored ? "colored" : "atomic";
    printf("=== Assembly-Free FEM Heat Solver v2 (Backward Euler + Jacobi-PCG) ===\n");
    printf("Grid %d^3 (DOFs %ld)  ranks %d (1D x-decomp, %d cells/rank)\n", o.N,
           (long)(o.N + 1) * (o.N + 1) * (o.N + 1), nprocs, o.N / nprocs);
    printf("dt %.4e  t_final %.4e  steps %d  kernel %s  scatter %s  mode %s\n", o.dt, o.t_final, n_steps,
           kname, sname, o.verify ? "VERIFY" : "production");
    if (!o.campaign.empty())
        printf("Campaign: %zu fault runs from %s (first in %s)\n", o.campaign_runs.size(), o.campaign.c_str(),
               o.campaign_runs.empty() ? "-" : o.campaign_runs[0].dir.c_str());
    std::vector<double> hk, hr;
    if (!load_fields(o, hk, hr)) return 1;

    Solver S;
    S.o = o;
    heat_config cfg;
    memset(&cfg, 0, sizeof cfg);
    cfg.n = o.N;
    cfg.dt = o.dt;
    cfg.kernel = o.kernel;
    cfg.scatter = o.colored ? HEAT_SCATTER_COLORED : HEAT_SCATTER_ATOMIC;
    cfg.verify = o.verify;
    cfg.no_source = o.no_source;
    cfg.source_mode = o.source_mode;
    cfg.source_value = o.source_value;
    cfg.detect = o.detect;
    const int rc = heat_setup(&cfg, hk.data(), hr.data(), &S.op);
    if (rc != HEAT_OK) {
        fprintf(stderr, "%s\n", heat_last_error());
        return rc == HEAT_E_SELFCHECK ? 2 : 1;
    }
    printf("Reference matrices: closed-form checks -> OK\n");
    int64_t nn = 0;
    API(heat_sizes(S.op, &nn, nullptr));
    S.nn = (long)nn;
    API(heat_field_stats(S.op, S.stats));
    API(heat_abft_threshold(S.op, &S.chk_tol));
    if (o.colored_impl >= 0) API(heat_set_colored_impl(S.op, o.colored_impl));
    API(heat_set_pcg_mode(S.op, o.pcg_order));
    API(heat_set_tuning(S.op, o.pcg_chunk, o.use_graph));
    API(heat_vector(S.op, HEAT_VEC_T, &S.T));
    API(heat_vector(S.op, HEAT_VEC_B, &S.b));
    API(heat_vector(S.op, HEAT_VEC_R, &S.r));
    API(heat_vector(S.op, HEAT_VEC_AP, &S.Ap));
    const double kmin = S.stats[0], kmax = S.stats[1], rmin = S.stats[2], rmax = S.stats[3];
    printf("Material: k in [%g, %g] (contrast %.3g), rho*c in [%g, %g]\n", kmin, kmax, kmax / kmin, rmin, rmax);

    int status = 0;
    if (o.verify) {
        // run_verify, src/heat_solver_het.cu:1293-1338
        if (kmin != kmax || rmin != rmax) {
            fprintf(stderr, "--verify needs a homogeneous field (k %g..%g, rhoc %g..%g)\n", kmin, kmax, rmin, rmax);
            heat_cleanup(S.op);
            return 1;
        }
        const double alpha = kmin / rmin;
        API(heat_initial_condition(S.op, S.T, 2, 0.0, alpha));
        API(heat_synchronize(S.op));
        const double t0 = now_ms();
        long total_it = 0;
        for (int step = 0; step < n_steps; ++step) {
            API(heat_build_rhs(S.op, S.T, S.b));
            total_it += S.pcg(S.T, S.b);
        }
        API(heat_synchronize(S.op));
        const double ms = now_ms() - t0;
        const double t_end = n_steps * o.dt;
        API(heat_initial_condition(S.op, S.r, 2, t_end, alpha));
        const double exact_norm = S.norm2(S.r);
        API(heat_axpy(S.op, -1.0, S.T, S.r));
        const double rel_err = S.norm2(S.r) / exact_norm;
        printf("VERIFY: N=%d dt=%.4e t_end=%.6e alpha=%g rel_L2_err=%.6e  (%s/%s, %.1f ms, %.1f CG it/step)\n",
               o.N, o.dt, t_end, alpha, rel_err, kname, sname, ms, (double)total_it / n_steps);
        const std::string p = o.outdir + "/verify.csv";
        FILE* vf = fopen(p.c_str(), "r");
        const bool fresh = (vf == nullptr);
        if (vf) fclose(vf);
        vf = fopen(p.c_str(), "a");
        if (!vf) {
            fprintf(stderr, "cannot write %s\n", p.c_str());
            heat_cleanup(S.op);
            return 1;
        }
        if (fresh) fprintf(vf, "N,dt,t_end,nprocs,kernel,scatter,rel_L2_err,solve_ms,avg_cg_iters\n");
        fprintf(vf, "%d,%.6e,%.6e,%d,%s,%s,%.10e,%.3f,%.3f\n", o.N, o.dt, t_end, nprocs, kname, sname, rel_err, ms,
                (double)total_it / n_steps);
        fclose(vf);
    } else if (!o.diag_cg.empty()) {
        FILE* lf = fopen(o.diag_cg.c_str(), "w");
        if (!lf) {
            fprintf(stderr, "cannot write %s\n", o.diag_cg.c_str());
            heat_cleanup(S.op);
            return 1;
        }
        fprintf(lf, "iter,rel_residual,alpha,beta\n");
        API(heat_initial_condition(S.op, S.T, o.ic_mode, 0.0, kmin / rmin));
        API(heat_build_rhs(S.op, S.T, S.b));
        std::vector<double> hist(3 * (size_t)(o.max_iter + 1));
        int it = 0;
        API(heat_pcg_solve(S.op, S.b, S.T, o.max_iter, o.tol, &it, hist.data()));
        fprintf(lf, "0,%.10e,,\n", hist[0]);
        for (int i = 1; i <= it; ++i)
            fprintf(lf, "%d,%.10e,%.10e,%.10e\n", i, hist[3 * i], hist[3 * i + 1], hist[3 * i + 2]);
        fclose(lf);
        printf("diag-cg: %d iterations to %g, contrast k_max/k_min=%.3g -> %s\n", it, o.tol, kmax / kmin,
               o.diag_cg.c_str());
    } else {
        // twin-run protocol, src/heat_solver_het.cu:1505-1621
        std::vector<std::vector<double>> snapA;
        RunResult rA, rB, rF;
        const CaseLabels lab = {n_steps, nprocs, kname, sname, kmin, kmax, rmin, rmax};
        run_case(S, n_steps, 0, -1, 0, 0, &snapA, nullptr, rA, "cleanA", nullptr);
        if (o.campaign.empty() && !write_field(snapA.back(), o.outdir, o.tag + "_clean")) status = 1;
        run_case(S, n_steps, 0, -1, 0, 0, nullptr, &snapA, rB, "cleanB", &rA);
        int on_bnd = -1;
        if (!o.campaign.empty()) {
            // the batch runs the fused FP64 colour-ordered solve; other solves
            // take the runs one after another
            const bool batched = o.campaign_batch > 0 && o.colored && o.kernel == HEAT_KERNEL_FAST && !o.mixed &&
                                 o.pcg_order == HEAT_PCG_FUSED;
            status = batched ? run_campaign_batched(S, cfg, hk, hr, n_steps, lab, snapA, rA, rB)
                             : run_campaign(S, n_steps, lab, snapA, rA, rB);
        } else if (o.fi_level > 0 && !status) {
            on_bnd = target_on_boundary(o);
            run_case(S, n_steps, o.fi_level, o.fi_step, o.fi_target, o.fi_bit, nullptr, &snapA, rF, "fault", &rA);
            std::vector<double> Tf;
            S.snapshot(Tf);
            if (!write_field(Tf, o.outdir, o.tag + "_fault")) status = 1;
        }
        if (!status && o.campaign.empty()) {
            if (!write_case(o, o.outdir, lab, rA, rB, rF, on_bnd)) status = 1;
        }
    }
    heat_cleanup(S.op);
    return status;
}

}  // namespace

int main(int argc, char** argv) {
    Opts o;
    const int prc = parse(argc, argv, o);
    if (prc) return prc;
    try {
        return run(o);
    } catch (const ApiError& e) {
        return 1;
    }
}
