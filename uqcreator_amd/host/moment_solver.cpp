// moment_solver.cpp -- MomentSolver on top of the C ABI: the reference's
// MomentSolver::Solve (momentsolver.cpp:37-269) with the closure, the sweeps
// and the time loop running in libuqk on the device.
//
//   constructor      Closure::Create tables + uqk_create        (momentsolver.cpp:4-30)
//   setup_ic         SetupIC: u_j = IC(x_j, xi_k) phiw           (momentsolver.cpp:535-563)
//   set_duals        SetDuals: lambda = (DS(u_0), 0, ...)         (momentsolver.cpp:463-523)
//   solve            the time loop                               (momentsolver.cpp:79-181)
//   mean_var         WriteMeanAndVar                             (momentsolver.cpp:302-328)
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <thread>
#include <vector>

#include "config.h"
#include "io.h"
#include "uqk.h"
#include "uqk_host.h"

namespace uqk_host {

namespace {

double closure_alpha(int kind) { return kind == UQK_EULER1D ? 0.5 : 1.0; }

// DS closed forms (EulerClosure2D::DS eulerclosure2d.cpp:90-102,
// ShallowWaterClosure2D::DS shallowwaterclosure2d.cpp:54-58, and the barrier
// variant); BoundedBarrier has none and leaves lambda at zero (closure.cpp:289).
void ds(const Settings& s, const double* u, double* out) {
    for (int l = 0; l < s.n_states; ++l) out[l] = 0.0;
    if (s.closure == UQK_EULER2D) {
        const double g = -s.gamma;
        const double rho = u[0], ru = u[1], rv = u[2], rE = u[3], q2 = ru * ru + rv * rv;
        out[0] = (q2 + g * (2 * rho * rE - q2)) / (-2 * rho * rE + q2) - std::log(std::pow(rho, g) * (rE - q2 / (2 * rho)));
        out[1] = -((2 * rho * ru) / (-2 * rho * rE + q2));
        out[2] = -((2 * rho * rv) / (-2 * rho * rE + q2));
        out[3] = -(rho / (rE - q2 / (2 * rho)));
    } else if (s.closure == UQK_SHALLOW_WATER2D) {
        out[0] = (9.81 * u[0] * u[0] * u[0] - 0.5 * u[1] * u[1] - 0.5 * u[2] * u[2]) / (u[0] * u[0]);
        out[1] = u[1] / u[0];
        out[2] = u[2] / u[0];
    } else if (s.closure == UQK_SHALLOW_WATER2D_BARRIER) {
        const double h = u[0], vx = u[1] / h, vy = u[2] / h;
        out[0] = 9.81 * h - s.barrier / h - 0.5 * (vx * vx + vy * vy);
        out[1] = vx;
        out[2] = vy;
    } else if (s.closure == UQK_EULER1D) {
        const double g = -s.gamma;
        const double rho = u[0], ru = u[1], ru2 = ru * ru, rE = u[2];
        out[0] = (ru2 + g * (2 * rho * rE - ru2)) / (-2 * rho * rE + ru2) - std::log(std::pow(rho, g) * (rE - ru2 / (2 * rho)));
        out[1] = -((2 * rho * ru) / (-2 * rho * rE + ru2));
        out[2] = -(rho / (rE - ru2 / (2 * rho)));
    }
}

template <class F>
void parallel_for(long n, F f) {
    const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nth; ++t)
        pool.emplace_back([&, t]() {
            for (long j = t; j < n; j += nth) f(j);
        });
    for (auto& th : pool) th.join();
}

struct DeviceBuf {
    void* p = nullptr;
    explicit DeviceBuf(size_t bytes) {
        p = uqk_device_alloc((int64_t)bytes);
        if (!p) throw std::runtime_error("device allocation failed");
    }
    DeviceBuf(const void* host, size_t bytes) : DeviceBuf(bytes) {
        if (uqk_memcpy(p, host, (int64_t)bytes, 0) != 0) throw std::runtime_error("host to device copy failed");
    }
    ~DeviceBuf() { uqk_device_free(p); }
    DeviceBuf(const DeviceBuf&) = delete;
    DeviceBuf& operator=(const DeviceBuf&) = delete;
    template <class T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

}  // namespace

struct Basis {
    std::vector<std::vector<int>> idx;
    std::vector<int> ntotal_per_level;
    std::vector<double> xi, wq, phi, phiw;
    int M = 0, Q = 0;
};

static Basis make_basis(const Settings& s) {
    Basis b;
    b.idx = index_set(s.dim_xi, s.degrees, s.max_degree, b.ntotal_per_level);
    if (!basis_tables(s.dim_xi, s.quad_points, b.idx, b.xi, b.wq, b.phi, b.phiw))
        throw std::runtime_error("quadrature failed");
    b.M = (int)b.idx.size();
    b.Q = (int)b.wq.size();
    return b;
}

static std::vector<double> setup_ic(const Settings& s, const Mesh& m, const Basis& b) {
    const int S = s.n_states, M = b.M, Q = b.Q, d = s.dim_xi;
    std::vector<double> u((size_t)m.n_cells * S * M, 0.0);
    parallel_for(m.n_cells, [&](long j) {
        std::vector<double> uq((size_t)Q * S);
        for (int k = 0; k < Q; ++k) initial_state(s, &m.center[(size_t)j * m.dim], &b.xi[(size_t)k * d], &uq[(size_t)k * S]);
        double* uj = &u[(size_t)j * S * M];
        for (int i = 0; i < M; ++i)
            for (int l = 0; l < S; ++l) {
                double acc = 0.0;
                for (int k = 0; k < Q; ++k) acc += uq[(size_t)k * S + l] * b.phiw[(size_t)k * M + i];
                uj[i * S + l] = acc;
            }
    });
    return u;
}

static std::vector<double> set_duals(const Settings& s, const std::vector<double>& u, long n, int M) {
    const int S = s.n_states;
    std::vector<double> lam(u.size(), 0.0);
    for (long j = 0; j < n; ++j) {
        double u0[4], d0[4];
        for (int l = 0; l < S; ++l) u0[l] = u[(size_t)j * S * M + l];
        ds(s, u0, d0);
        for (int l = 0; l < S; ++l) lam[(size_t)j * S * M + l] = d0[l];
    }
    return lam;
}

static uqk_ctx* make_ctx(const Settings& s, const Basis& b, int max_iterations, bool safe) {
    uqk_desc d{};
    d.closure = s.closure;
    d.n_states = s.n_states;
    d.n_moments = b.M;
    d.n_quad = b.Q;
    d.dim_xi = s.dim_xi;
    d.gamma = s.gamma;
    d.u_minus = s.u_minus;
    d.u_plus = s.u_plus;
    d.gravity = 9.81;
    d.barrier = s.barrier;
    d.alpha = closure_alpha(s.closure);
    d.epsilon = s.epsilon;
    d.max_iterations = max_iterations;
    d.max_halvings = 1000;
    d.du_exact = s.du_exact ? 1 : 0;
    d.safe_line_search = safe ? 1 : 0;
    d.precision = s.precision;
    d.phi = b.phi.data();
    d.phiw = b.phiw.data();
    d.xi = b.xi.data();
    d.device = 0;
    uqk_ctx* ctx = nullptr;
    if (uqk_create(&d, &ctx) != 0) {
        std::string msg = ctx ? uqk_last_error(ctx) : "uqk_create failed";
        uqk_destroy(ctx);
        throw std::runtime_error(msg);
    }
    return ctx;
}

struct RunOutput {
    long n_cells = 0;
    int64_t steps = 0, newton = 0;
    double t = 0.0, residual = 0.0, runtime_s = 0.0;
    int S = 0, M = 0, Q = 0;
};

// The start of a run: moments and duals, the start time, and whether the
// duals must be converged first.  Without a restart file: SetupIC and
// SetDuals' (DS(mean), 0, ...).  With one (DetermineMoments,
// DeterminePreviousSettings, SetDuals; momentsolver.cpp:418-523): the
// moments of <restartFile>_moments, the duals of <restartFile>_duals when
// importDualState is set, and the previous run's tEnd as the start time.
// When the previous run used other moments or another quadrature,
// restart_other_basis below maps its state onto the current basis.
struct Start {
    std::vector<double> u, lam;
    double t0 = 0.0;
    bool presolve = false;
};

// A restart whose previous run used other moments or another quadrature
// (DetermineMoments, SetDuals; momentsolver.cpp:418-523), on the GPU:
//  1. the previous moments (its M_p per cell); when M_p differs from M,
//     Dirichlet cells take the initial condition at the current order, of
//     which the previous closure reads the first M_p moments (:423-429);
//  2. the previous duals when importDualState is set, else (DS(u_0), 0, ...)
//     converged by the previous closure (its basis and quadrature, 10000
//     iterations, SolveClosureSafe) when the moments changed or the closure is
//     one-shot (:466-493);
//  3. when the moments changed (:495-521): Lambda of the previous basis at
//     the current quadrature points, U there, projected on the current basis
//     gives the moments; the duals are the previous ones padded with zeros
//     (or cut) to M, converged by the current closure (SolveClosureSafe;
//     10000 iterations for a one-shot closure).
// The start-up closures run in FP64 whatever the run's precision.
static void restart_other_basis(const Settings& s, const Settings& prev, const Mesh& m, const Basis& b,
                                const std::string& rf, Start& st) {
    const long N = m.n_cells;
    const int S = s.n_states, M = b.M;
    Settings ps = s;  // DeterminePreviousSettings: the previous basis, the current closure
    ps.degrees = prev.degrees;
    ps.max_degree = prev.max_degree;
    ps.quad_points = prev.quad_points;
    ps.precision = UQK_FP64;
    Settings cs = s;
    cs.precision = UQK_FP64;
    const Basis bp = make_basis(ps);
    const int Mp = bp.M;
    const bool moments_changed = prev.degrees != s.degrees || prev.max_degree != s.max_degree;
    std::vector<double> u = read_restart_field(rf + "_moments", N, S, Mp);
    if (Mp != M) {
        const std::vector<double> uic = setup_ic(s, m, b);
        for (long j = 0; j < N; ++j)
            if (m.bc[j] == UQK_BC_DIRICHLET)
                for (int i = 0; i < std::min(Mp, M); ++i)
                    for (int l = 0; l < S; ++l) u[(size_t)j * S * Mp + i * S + l] = uic[(size_t)j * S * M + i * S + l];
    }
    std::vector<double> lam;
    if (s.import_dual_state) {
        lam = read_restart_field(rf + "_duals", N, S, Mp);
    } else {
        lam = set_duals(s, u, N, Mp);
        if (s.max_iterations == 1 || moments_changed) {
            std::unique_ptr<uqk_ctx, int (*)(uqk_ctx*)> cp(make_ctx(ps, bp, 10000, true), uqk_destroy);
            DeviceBuf du(u.data(), sizeof(double) * u.size()), dl(lam.data(), sizeof(double) * lam.size());
            if (uqk_closure_solve(cp.get(), N, nullptr, du.as<double>(), dl.as<double>(), nullptr, nullptr) < 0)
                throw std::runtime_error(uqk_last_error(cp.get()));
            if (uqk_memcpy(lam.data(), dl.p, (int64_t)(sizeof(double) * lam.size()), 1) != 0)
                throw std::runtime_error("device to host copy failed");
        }
    }
    if (!moments_changed) {  // another quadrature only: the same moments and duals
        st.u = std::move(u);
        st.lam = std::move(lam);
        return;
    }
    Basis bi;  // the previous index set on the current quadrature
    bi.idx = bp.idx;
    if (!basis_tables(s.dim_xi, s.quad_points, bi.idx, bi.xi, bi.wq, bi.phi, bi.phiw))
        throw std::runtime_error("quadrature failed");
    bi.M = Mp;
    bi.Q = b.Q;
    std::unique_ptr<uqk_ctx, int (*)(uqk_ctx*)> ci(make_ctx(cs, bi, 1, true), uqk_destroy);
    std::unique_ptr<uqk_ctx, int (*)(uqk_ctx*)> cn(
        make_ctx(cs, b, s.max_iterations == 1 ? 10000 : s.max_iterations, true), uqk_destroy);
    DeviceBuf dlp(lam.data(), sizeof(double) * lam.size());
    DeviceBuf duq(sizeof(double) * (size_t)N * b.Q * S), dun(sizeof(double) * (size_t)N * M * S);
    if (uqk_eval_uq(ci.get(), N, nullptr, dlp.as<double>(), duq.as<double>()) != 0)
        throw std::runtime_error(uqk_last_error(ci.get()));
    if (uqk_reconstruct(cn.get(), N, nullptr, duq.as<double>(), dun.as<double>()) != 0)
        throw std::runtime_error(uqk_last_error(cn.get()));
    std::vector<double> ln((size_t)N * M * S, 0.0);
    for (long j = 0; j < N; ++j)
        for (int i = 0; i < std::min(Mp, M); ++i)
            for (int l = 0; l < S; ++l) ln[(size_t)j * S * M + i * S + l] = lam[(size_t)j * S * Mp + i * S + l];
    DeviceBuf dln(ln.data(), sizeof(double) * ln.size());
    if (uqk_closure_solve(cn.get(), N, nullptr, dun.as<double>(), dln.as<double>(), nullptr, nullptr) < 0)
        throw std::runtime_error(uqk_last_error(cn.get()));
    st.u.assign((size_t)N * M * S, 0.0);
    if (uqk_memcpy(st.u.data(), dun.p, (int64_t)(sizeof(double) * st.u.size()), 1) != 0 ||
        uqk_memcpy(ln.data(), dln.p, (int64_t)(sizeof(double) * ln.size()), 1) != 0)
        throw std::runtime_error("device to host copy failed");
    st.lam = std::move(ln);
}

static Start start_of(const Settings& s, const Mesh& m, const Basis& b) {
    Start st;
    const long N = m.n_cells;
    const int S = s.n_states, M = b.M;
    if (s.restart_file.empty()) {
        st.u = setup_ic(s, m, b);
        st.lam = set_duals(s, st.u, N, M);
        st.presolve = s.max_iterations == 1;
        return st;
    }
    const std::string rf = s.restart_file[0] == '/' ? s.restart_file : s.input_dir + "/" + s.restart_file;
    const Settings prev = Settings::from_toml(Toml::parse(read_log_config(rf)));
    if (prev.dim_xi != s.dim_xi)
        throw std::runtime_error("restart: the previous run had another number of uncertain dimensions");
    if (prev.degrees != s.degrees || prev.max_degree != s.max_degree || prev.quad_points != s.quad_points) {
        restart_other_basis(s, prev, m, b, rf, st);
        st.t0 = prev.t_end;
        return st;
    }
    st.u = read_restart_field(rf + "_moments", N, S, M);
    st.lam = s.import_dual_state ? read_restart_field(rf + "_duals", N, S, M) : set_duals(s, st.u, N, M);
    st.t0 = prev.t_end;
    st.presolve = s.max_iterations == 1 && !s.import_dual_state;
    return st;
}

static RunOutput run_config(const std::string& path, const char* mean_var_csv, const char* output_prefix) {
    const Settings s = Settings::from_file(path);
    const Basis b = make_basis(s);
    const Mesh m = Mesh::from_settings(s);
    const int S = s.n_states, M = b.M, Q = b.Q;
    const long N = m.n_cells;
    Start start = start_of(s, m, b);
    std::vector<double>& u = start.u;
    std::vector<double>& lam = start.lam;

    RunOutput out;
    out.n_cells = N;
    out.S = S;
    out.M = M;
    out.Q = Q;
    DeviceBuf du(u.data(), sizeof(double) * u.size()), dl(lam.data(), sizeof(double) * lam.size());
    DeviceBuf duq(sizeof(double) * (size_t)N * Q * S), dus(sizeof(double) * u.size());
    DeviceBuf dst(sizeof(int32_t) * N), dit(sizeof(int32_t) * N);
    DeviceBuf nb(m.nbr.data(), sizeof(int32_t) * m.nbr.size()), un(m.unit_normal.data(), sizeof(double) * m.unit_normal.size()),
        sn(m.normal.data(), sizeof(double) * m.normal.size()), ar(m.area.data(), sizeof(double) * m.area.size()),
        dx(m.dx_param.data(), sizeof(double) * m.dx_param.size()), bc(m.bc.data(), sizeof(int32_t) * m.bc.size()),
        bn(m.bnormal.data(), sizeof(double) * m.bnormal.size());
    uqk_mesh mesh{N, m.n_edges, m.dim, nb.as<int32_t>(), un.as<double>(), sn.as<double>(), ar.as<double>(),
                  dx.as<double>(), bc.as<int32_t>(), bn.as<double>()};
    uqk_problem prob{};
    prob.problem = s.problem == "Burgers1D" ? UQK_BURGERS : s.problem == "Euler2D" ? UQK_EULER2D_PROBLEM
                                                                                     : UQK_SHALLOW_WATER2D_PROBLEM;
    prob.flux = s.flux;
    prob.cfl = s.cfl;
    prob.gamma = s.gamma;
    prob.gravity = 9.81;
    prob.friction = s.friction;
    prob.friction_dim = s.friction_dim;

    if (start.presolve) {
        // one-shot IPM starts from converged duals (momentsolver.cpp:487-493)
        uqk_ctx* c0 = make_ctx(s, b, 10000, true);
        int64_t f = uqk_closure_solve(c0, N, nullptr, du.as<double>(), dl.as<double>(), nullptr, nullptr);
        uqk_destroy(c0);
        if (f != 0) throw std::runtime_error("initial closure failed");
    }
    std::unique_ptr<uqk_ctx, int (*)(uqk_ctx*)> owner(make_ctx(s, b, s.max_iterations, s.safe_line_search),
                                                      uqk_destroy);
    uqk_ctx* ctx = owner.get();
    // refinement levels (momentsolver.cpp:39): every cell starts on the finest
    const int nlev = (int)b.ntotal_per_level.size();
    std::vector<int32_t> nt_host;
    if (nlev > 1) {
        nt_host.assign((size_t)N, b.ntotal_per_level.back());
        if (uqk_set_refinement(ctx, nlev, b.ntotal_per_level.data(), s.refine_above, s.coarsen_below) != 0)
            throw std::runtime_error(uqk_last_error(ctx));
    }
    std::unique_ptr<DeviceBuf> dnt;
    if (nlev > 1) dnt.reset(new DeviceBuf(nt_host.data(), sizeof(int32_t) * nt_host.size()));
    int32_t* nt = dnt ? dnt->as<int32_t>() : nullptr;
    auto tic = std::chrono::steady_clock::now();
    int64_t nfail = 0;
    if (s.min_residual > 0.0) {
        // steady runs stop on the residual (momentsolver.cpp:79): step by step.
        // With refinement retardation (Paper1/adaptivity.tex:61-86) the cells
        // start on level 0 under a level cap that rises by one whenever the
        // residual falls below the next tolerance; the run ends on the
        // residual only once the cap has reached the finest level.
        const bool retard = nlev > 1 && !s.retardation.empty();
        size_t rl = 0;
        auto set_cap = [&](int cap) {
            if (uqk_set_refinement_cap(ctx, cap) != 0) throw std::runtime_error(uqk_last_error(ctx));
        };
        if (retard) {
            set_cap(0);
            std::vector<int32_t> nt0((size_t)N, b.ntotal_per_level[0]);
            if (uqk_memcpy(nt, nt0.data(), (int64_t)(sizeof(int32_t) * nt0.size()), 0) != 0)
                throw std::runtime_error("host to device copy failed");
        }
        double t = start.t0, res = s.min_residual + 1.0;
        while (t < s.t_end && (res > s.min_residual || (retard && rl < s.retardation.size())) &&
               (s.max_steps < 0 || out.steps < s.max_steps)) {
            double dt = 0.0;
            nfail = uqk_step(ctx, &prob, &mesh, nt, du.as<double>(), dl.as<double>(), duq.as<double>(),
                             dus.as<double>(), dst.as<int32_t>(), dit.as<int32_t>(), &dt, &res);
            if (nfail) break;
            int64_t a = 0, g = 0;
            uqk_last_closure_stats(ctx, &a, &g);
            out.newton += a;
            t += dt;
            ++out.steps;
            if (retard) {
                const size_t rl0 = rl;
                while (rl < s.retardation.size() && res < s.retardation[rl]) ++rl;
                if (rl != rl0) set_cap((int)std::min<size_t>(rl, (size_t)nlev - 1));
            }
        }
        out.t = t;
        out.residual = res;
    } else {
        nfail = uqk_run_steps(ctx, &prob, &mesh, nt, du.as<double>(), dl.as<double>(), duq.as<double>(),
                              dus.as<double>(), dst.as<int32_t>(), dit.as<int32_t>(), start.t0, s.t_end, s.max_steps, 1,
                              &out.t, &out.steps, &out.residual, &out.newton);
    }
    auto toc = std::chrono::steady_clock::now();
    out.runtime_s = std::chrono::duration<double>(toc - tic).count();
    if (nfail < 0) throw std::runtime_error(uqk_last_error(ctx));
    if (nfail > 0) throw std::runtime_error("[closure] Newton failed in " + std::to_string(nfail) + " cells");
    if (!mean_var_csv && !output_prefix) return out;
    std::vector<double> h((size_t)N * 2 * S);
    {
        DeviceBuf mv(sizeof(double) * h.size());
        if (uqk_mean_var(ctx, N, nt, dl.as<double>(), mv.as<double>()) != 0) throw std::runtime_error(uqk_last_error(ctx));
        if (uqk_memcpy(h.data(), mv.p, (int64_t)(sizeof(double) * h.size()), 1) != 0)
            throw std::runtime_error("device to host copy failed");
    }
    if (mean_var_csv) {
        std::ofstream f(mean_var_csv);
        f.precision(17);
        for (long j = 0; j < N; ++j) {
            for (int q = 0; q < 2 * S; ++q) f << (q ? "," : "") << h[(size_t)j * 2 * S + q];
            f << "\n";
        }
        if (!f) throw std::runtime_error(std::string("write failed: ") + mean_var_csv);
    }
    if (output_prefix) {
        // the end of MomentSolver::Solve (momentsolver.cpp:221-228): fields,
        // then the moments and duals a restart reads beside the log
        const std::string pre = output_prefix;
        write_fields(pre, s, m, h);
        if (uqk_memcpy(u.data(), du.p, (int64_t)(sizeof(double) * u.size()), 1) != 0 ||
            uqk_memcpy(lam.data(), dl.p, (int64_t)(sizeof(double) * lam.size()), 1) != 0)
            throw std::runtime_error("device to host copy failed");
        char line[160];
        std::snprintf(line, sizeof line, "steps %lld  t %.17g  residual %.6e  runtime %.3f s", (long long)out.steps,
                      out.t, out.residual, out.runtime_s);
        write_log(pre, s.input_path, s.input_text, {line});
        write_restart_field(pre + "_moments", u, N, S, M);
        write_restart_field(pre + "_duals", lam, N, S, M);
    }
    return out;
}

}  // namespace uqk_host

extern "C" {

int uqk_run_config(const char* toml_path, const char* mean_var_csv, uqk_run_report* report, char* err,
                   int32_t errlen) {
    return uqk_run_config_output(toml_path, mean_var_csv, nullptr, report, err, errlen);
}

int uqk_run_config_output(const char* toml_path, const char* mean_var_csv, const char* output_prefix,
                          uqk_run_report* report, char* err, int32_t errlen) {
    try {
        uqk_host::RunOutput o = uqk_host::run_config(toml_path, mean_var_csv, output_prefix);
        if (report) {
            report->n_cells = o.n_cells;
            report->steps = o.steps;
            report->t = o.t;
            report->residual = o.residual;
            report->runtime_s = o.runtime_s;
            report->cell_steps_per_s = o.runtime_s > 0 ? (double)o.n_cells * (double)o.steps / o.runtime_s : 0.0;
            report->newton_steps = o.newton;
            report->n_states = o.S;
            report->n_moments = o.M;
            report->n_quad = o.Q;
        }
        return 0;
    } catch (const std::exception& e) {
        if (err && errlen > 0) {
            std::strncpy(err, e.what(), (size_t)errlen - 1);
            err[errlen - 1] = '\0';
        }
        return -1;
    }
}

int64_t uqk_config_setup(const char* toml_path, int32_t* info, double* u, double* lambda, double* center, char* err,
                         int32_t errlen) {
    try {
        const uqk_host::Settings s = uqk_host::Settings::from_file(toml_path);
        const uqk_host::Basis b = uqk_host::make_basis(s);
        const uqk_host::Mesh m = uqk_host::Mesh::from_settings(s);
        if (info) {
            info[0] = s.closure;
            info[1] = s.n_states;
            info[2] = b.M;
            info[3] = b.Q;
            info[4] = s.dim_xi;
            info[5] = m.n_edges;
            info[6] = m.dim;
            info[7] = s.precision;
        }
        if (u || lambda) {
            const uqk_host::Start st = uqk_host::start_of(s, m, b);
            if (u) std::copy(st.u.begin(), st.u.end(), u);
            if (lambda) std::copy(st.lam.begin(), st.lam.end(), lambda);
        }
        if (center) std::copy(m.center.begin(), m.center.end(), center);
        return m.n_cells;
    } catch (const std::exception& e) {
        if (err && errlen > 0) {
            std::strncpy(err, e.what(), (size_t)errlen - 1);
            err[errlen - 1] = '\0';
        }
        return -1;
    }
}

}  // extern "C"
