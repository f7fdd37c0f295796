// Physics diagnostics and the headless pd_bench-style report.
//
// physicsDiagnostics restates the reference's end-of-run figures in its own
// summation order, so the printed %.6g values can be compared digit for digit:
//   Mesh::maxRelativeStrain / meanRelativeStrain / totalMass  core/mesh/Mesh.cpp:345-372
//   surrogateEnergy / totalEnergy / maxSpeed                   core/sim/Integrator.cpp:804-846
//
// pdx_bench_run is a caller of the public C ABI, in the role pd_bench
// (app/bench/main.cpp:53-264) plays for pd::stepOnce: the same options and
// defaults, the scene makeScene builds (core/sim/Scene.cpp:39-169), one substep
// per call with the strain tracked after each, and the same report sections
// (setup, results, stage times, factor reuse).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pdx.h"
#include "host.h"

namespace pdx {

void physicsDiagnostics(const Scene& s, const double* x, const double* v, const double g[3], double out[6]) {
  auto edgeLength = [&](int e) {
    const int a = s.ea[(size_t)e], b = s.eb[(size_t)e];
    const double dx = x[3 * a] - x[3 * b], dy = x[3 * a + 1] - x[3 * b + 1], dz = x[3 * a + 2] - x[3 * b + 2];
    return std::sqrt(dx * dx + dy * dy + dz * dz);
  };
  double worst = 0.0, sum = 0.0, elastic = 0.0;
  for (int e = 0; e < s.E; ++e) {
    const double len = edgeLength(e), rest = s.rest[(size_t)e];
    if (rest <= 0.0) continue;
    const double strain = std::fabs(len - rest) / rest;
    worst = std::max(worst, strain);
    sum += strain;
  }
  for (int e = 0; e < s.E; ++e) {
    const double d = edgeLength(e) - s.rest[(size_t)e];
    elastic += 0.5 * s.kappa[(size_t)e] * d * d;
  }
  for (int q = 0; q < s.B; ++q) {
    // second difference written a - 2b + c, as the bend force is
    const double* pa = x + 3 * s.sa[(size_t)q];
    const double* pb = x + 3 * s.sb[(size_t)q];
    const double* pc = x + 3 * s.sc[(size_t)q];
    const double sx = pa[0] - 2.0 * pb[0] + pc[0], sy = pa[1] - 2.0 * pb[1] + pc[1],
                 sz = pa[2] - 2.0 * pb[2] + pc[2];
    elastic += 0.5 * s.bk[(size_t)q] * (sx * sx + sy * sy + sz * sz);
  }
  double kinetic = 0.0, potential = 0.0, fastest = 0.0, mass = 0.0;
  for (int i = 0; i < s.n; ++i) {
    const double m = s.mass[(size_t)i];
    const double* vi = v + 3 * i;
    const double* xi = x + 3 * i;
    const double v2 = vi[0] * vi[0] + vi[1] * vi[1] + vi[2] * vi[2];
    kinetic += 0.5 * m * v2;
    potential -= m * (g[0] * xi[0] + g[1] * xi[1] + g[2] * xi[2]);
    fastest = std::max(fastest, std::sqrt(v2));
    mass += m;
  }
  out[0] = worst;
  out[1] = s.E > 0 ? sum / static_cast<double>(s.E) : 0.0;
  out[2] = elastic;
  out[3] = kinetic + elastic + potential;
  out[4] = fastest;
  out[5] = mass;
}

}  // namespace pdx

namespace {

// Iteration presets (core/sim/Scene.cpp:213-226): name, iteration cap, residual tolerance.
struct Preset {
  const char* name;
  int iters;
  double residualTol;
};
const Preset kPresets[] = {{"preview", 20, 0.3}, {"realtime", 40, 0.3}, {"accurate", 80, 1.0e-3}};

struct Text {
  std::string s;
  void add(const char* fmt, ...) __attribute__((format(printf, 2, 3))) {
    va_list ap, again;
    va_start(ap, fmt);
    va_copy(again, ap);
    const int k = std::vsnprintf(nullptr, 0, fmt, ap);
    va_end(ap);
    if (k > 0) {
      std::vector<char> buf((size_t)k + 1);
      std::vsnprintf(buf.data(), buf.size(), fmt, again);
      s.append(buf.data(), (size_t)k);
    }
    va_end(again);
  }
};

const char* kUsage =
    "usage: pdx_bench [options]\n"
    "  --grid NX NY        regular grid (default 40 40)\n"
    "  --mesh PATH         load an OBJ triangle mesh (overrides --grid)\n"
    "  --spacing S         grid spacing (default 0.025)\n"
    "  --steps N           substeps (default 600)\n"
    "  --substeps N        substeps per frame (default 2; reported only)\n"
    "  --dt H              substep length (default 1/120)\n"
    "  --stiffness K       distance-constraint stiffness (default 1e4)\n"
    "  --density RHO       areal density (default 1.0)\n"
    "  --damping KD        velocity damping (default 0.02)\n"
    "  --iters N           PD iteration cap per substep (default 10)\n"
    "  --preset NAME       preview = 20 / 0.3|g|, realtime = 40 / 0.3|g|, accurate = 80 / 1e-3;\n"
    "                      explicit --iters / --residual-tol override it in any order\n"
    "  --residual-tol T    residual criterion: unbalanced force <= T x |g| (default 1e-3; <= 0 disables)\n"
    "  --tol T             displacement criterion relTolerance (default 1e-3; decides only when\n"
    "                      the residual criterion is disabled)\n"
    "  --threads N         accepted for compatibility; the GPU substep has no thread count\n"
    "  --pin-top           pin the two top corners (the default)\n"
    "  --pin-all           pin every vertex\n"
    "  --report N          print a status line every N substeps\n"
    "  --shear K           shear diagonals with stiffness K (grids; also PD_SHEAR)\n"
    "  --bend K            linear bend stencils with stiffness K (grids; also PD_BEND)\n"
    "  --precision P       fp64 (default, the drop-in) or mixed32\n"
    "  --solve S           exact (default, the reference's sum order), fast (chain-blocked\n"
    "                      sweeps) or dense (the inverse formed once, one product per iteration)\n"
    "  --cheb-rho R        Chebyshev blend with spectral radius R (also PD_CHEB=1, PD_CHEB_RHO)\n"
    "  --cheb-start S      plain iterations before the blend (default 10; also PD_CHEB_START)\n"
    "  --device D          HIP device (default 0)\n";

struct Options {
  int nx = 40, ny = 40;
  std::string mesh;
  double spacing = 0.025;
  int steps = 600, substeps = 2, threads = 0, reportEvery = 0, device = 0;
  double dt = 1.0 / 120.0, stiffness = 1.0e4, density = 1.0, damping = 0.02;
  int iters = 10;
  double residualTol = 1.0e-3, relTol = 1.0e-3, absTol = 1.0e-12;
  double shear = 0.0, bend = 0.0;
  double chebRho = 0.0;  // > 0: Chebyshev blend
  int chebStart = 10;
  bool pinAll = false, help = false;
  int preset = -1;
  int precision = PDX_PREC_FP64, solve = PDX_SOLVE_EXACT;
};

// Returns 0, or -1 with a message in `err`.
int parse(int argc, const char* const* argv, Options& o, std::string& err) {
  bool explicitIters = false, explicitResidual = false;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i] ? argv[i] : "";
    auto real = [&](double& out) {
      if (i + 1 < argc) out = std::atof(argv[++i]);
    };
    auto whole = [&](int& out) {
      if (i + 1 < argc) out = std::atoi(argv[++i]);
    };
    if (a == "--grid" && i + 2 < argc) {
      o.nx = std::atoi(argv[++i]);
      o.ny = std::atoi(argv[++i]);
    } else if (a == "--mesh" && i + 1 < argc) {
      o.mesh = argv[++i];
    } else if (a == "--spacing") {
      real(o.spacing);
    } else if (a == "--steps") {
      whole(o.steps);
    } else if (a == "--substeps") {
      whole(o.substeps);
    } else if (a == "--dt") {
      real(o.dt);
    } else if (a == "--stiffness") {
      real(o.stiffness);
    } else if (a == "--density") {
      real(o.density);
    } else if (a == "--damping") {
      real(o.damping);
    } else if (a == "--iters") {
      whole(o.iters);
      explicitIters = true;
    } else if (a == "--preset") {
      int found = -1;
      if (i + 1 < argc)
        for (int k = 0; k < 3; ++k)
          if (std::strcmp(argv[i + 1], kPresets[k].name) == 0) found = k;
      if (found < 0) {
        err = "--preset takes one of: preview, realtime, accurate\n";
        return -1;
      }
      ++i;
      o.preset = found;
    } else if (a == "--residual-tol") {
      real(o.residualTol);
      explicitResidual = true;
    } else if (a == "--tol") {
      real(o.relTol);
    } else if (a == "--threads") {
      whole(o.threads);
    } else if (a == "--pin-all") {
      o.pinAll = true;
    } else if (a == "--report") {
      whole(o.reportEvery);
    } else if (a == "--shear") {
      real(o.shear);
    } else if (a == "--bend") {
      real(o.bend);
    } else if (a == "--device") {
      whole(o.device);
    } else if (a == "--cheb-rho") {
      real(o.chebRho);
    } else if (a == "--cheb-start") {
      whole(o.chebStart);
    } else if (a == "--precision" && i + 1 < argc) {
      const std::string p = argv[++i];
      if (p != "fp64" && p != "mixed32") {
        err = "--precision takes fp64 or mixed32\n";
        return -1;
      }
      o.precision = p == "fp64" ? PDX_PREC_FP64 : PDX_PREC_MIXED32;
    } else if (a == "--solve" && i + 1 < argc) {
      const std::string p = argv[++i];
      if (p != "exact" && p != "fast" && p != "dense") {
        err = "--solve takes exact, fast or dense\n";
        return -1;
      }
      o.solve = p == "exact" ? PDX_SOLVE_EXACT : (p == "fast" ? PDX_SOLVE_FAST : PDX_SOLVE_DENSE);
    } else if (a == "--help" || a == "-h") {
      o.help = true;
    }
    // --pin-top (the default) and unknown options: nothing to do
  }
  if (o.preset >= 0) {
    // the preset first, then the explicit values, whatever the order on the line
    if (!explicitIters) o.iters = kPresets[o.preset].iters;
    if (!explicitResidual) o.residualTol = kPresets[o.preset].residualTol;
  }
  // shear / bend stiffness from the environment override the options when positive
  // (core/sim/Scene.cpp:80-85, :108-112)
  if (const char* env = std::getenv("PD_SHEAR")) {
    const double k = std::atof(env);
    if (k > 0.0) o.shear = k;
  }
  if (const char* env = std::getenv("PD_BEND")) {
    const double k = std::atof(env);
    if (k > 0.0) o.bend = k;
  }
  // the Chebyshev blend switches the reference's substep reads (Integrator.cpp:493-503):
  // PD_CHEB set, non-empty and not starting with '0' turns it on, rho 0.98 and
  // start 10 unless PD_CHEB_RHO / PD_CHEB_START give them
  if (const char* env = std::getenv("PD_CHEB")) {
    if (*env != '\0' && *env != '0') {
      const char* r = std::getenv("PD_CHEB_RHO");
      const char* s = std::getenv("PD_CHEB_START");
      o.chebRho = (r && *r) ? std::atof(r) : 0.98;
      o.chebStart = (s && *s) ? std::atoi(s) : 10;
    }
  }
  return 0;
}

// The cloth makeScene builds (core/sim/Scene.cpp:39-169).
struct Built {
  int n = 0, E = 0, B = 0;
  std::vector<double> x, mass;
  std::vector<uint8_t> pinned;
  std::vector<int32_t> ea, eb, sa, sb, sc;
  std::vector<double> rest, kap, bk;
  bool fromMesh = false;
};

int buildGrid(const Options& o, Built& s) {
  int n = 0, E = 0, B = 0, Es = 0, Bz = 0;
  int rc = pdx_grid_counts(o.nx, o.ny, o.shear > 0.0, o.bend > 0.0, &n, &E, &B);
  if (rc) return rc;
  pdx_grid_counts(o.nx, o.ny, 0, 0, nullptr, &Es, &Bz);
  s.n = n, s.E = E, s.B = B;
  s.x.resize(3 * (size_t)n);
  s.ea.resize(E), s.eb.resize(E), s.rest.resize(E), s.kap.resize(E);
  s.sa.resize(B), s.sb.resize(B), s.sc.resize(B), s.bk.resize(B);
  rc = pdx_grid_fill(o.nx, o.ny, o.spacing, o.spacing, o.stiffness, o.shear, o.bend, s.x.data(), s.ea.data(),
                     s.eb.data(), s.rest.data(), s.kap.data(), B ? s.sa.data() : nullptr, B ? s.sb.data() : nullptr,
                     B ? s.sc.data() : nullptr, B ? s.bk.data() : nullptr);
  if (rc) return rc;
  // masses from the structural edges, before the shear diagonals are appended
  s.mass.resize(n);
  rc = pdx_masses_from_edges(n, Es, s.ea.data(), s.eb.data(), s.rest.data(), o.density > 0.0 ? o.density : 1.0,
                             s.mass.data());
  if (rc) return rc;
  s.pinned.assign(n, 0);
  s.pinned[(size_t)(o.ny - 1) * o.nx] = 1;
  s.pinned[(size_t)(o.ny - 1) * o.nx + o.nx - 1] = 1;
  return PDX_OK;
}

int buildMesh(const Options& o, Built& s) {
  int n = 0, E = 0, T = 0;
  int rc = pdx_obj_counts(o.mesh.c_str(), &n, &E, &T);
  if (rc) return rc;
  std::vector<int32_t> tris(3 * (size_t)T);
  s.n = n, s.E = E, s.B = 0;
  s.x.resize(3 * (size_t)n);
  s.ea.resize(E), s.eb.resize(E), s.rest.resize(E), s.kap.resize(E);
  rc = pdx_obj_fill(o.mesh.c_str(), o.stiffness, s.x.data(), tris.data(), s.ea.data(), s.eb.data(), s.rest.data(),
                    s.kap.data());
  if (rc) return rc;
  s.mass.resize(n);
  rc = pdx_masses_from_surfaces(n, s.x.data(), T, tris.data(), o.density > 0.0 ? o.density : 1.0, s.mass.data());
  if (rc) return rc;
  s.pinned.assign(n, 0);  // a loaded mesh has no default pins
  s.fromMesh = true;
  return PDX_OK;
}

}  // namespace

extern "C" int pdx_bench_run(int argc, const char* const* argv, char* text, long long text_cap,
                             pdx_bench_summary* summary) {
  Text out;
  auto finish = [&](int rc) {
    if (text && text_cap > 0) {
      const size_t k = std::min<size_t>(out.s.size(), (size_t)(text_cap - 1));
      std::memcpy(text, out.s.data(), k);
      text[k] = '\0';
    }
    return rc;
  };
  Options o;
  std::string err;
  if (parse(argc, argv, o, err)) {
    out.add("%s", err.c_str());
    return finish(PDX_EINVAL);
  }
  if (o.help) {
    out.add("%s", kUsage);
    return finish(PDX_OK);
  }
  Built s;
  int rc = PDX_EINVAL;
  if (!o.mesh.empty()) {
    rc = buildMesh(o, s);
    if (rc) out.add("mesh %s did not load (status %d): using the %dx%d grid\n", o.mesh.c_str(), rc, o.nx, o.ny);
  }
  if (rc) {
    s = Built{};
    if ((rc = buildGrid(o, s))) {
      out.add("grid %dx%d rejected (status %d)\n", o.nx, o.ny, rc);
      return finish(rc);
    }
  }
  if (o.pinAll) std::fill(s.pinned.begin(), s.pinned.end(), 1);
  int npin = 0;
  for (uint8_t p : s.pinned) npin += p != 0;

  pdx_ctx* ctx = nullptr;
  if ((rc = pdx_create(&ctx, o.device, o.precision, o.solve))) {
    out.add("pdx_create failed (status %d)\n", rc);
    return finish(rc);
  }
  auto fail = [&](const char* what, int code) {
    out.add("%s failed (status %d): %s\n", what, code, pdx_last_error(ctx));
    pdx_destroy(ctx);
    return finish(code);
  };
  pdx_scene sc{};
  sc.n_vertices = s.n;
  sc.positions = s.x.data();
  sc.masses = s.mass.data();
  sc.pinned = s.pinned.data();
  sc.n_edges = s.E;
  sc.edge_a = s.ea.data();
  sc.edge_b = s.eb.data();
  sc.edge_rest = s.rest.data();
  sc.edge_kappa = s.kap.data();
  sc.n_bends = s.B;
  if (s.B) {
    sc.bend_a = s.sa.data();
    sc.bend_b = s.sb.data();
    sc.bend_c = s.sc.data();
    sc.bend_k = s.bk.data();
  }
  pdx_config cfg{};
  cfg.dt = o.dt;
  cfg.gravity[0] = 0.0, cfg.gravity[1] = -9.81, cfg.gravity[2] = 0.0;
  cfg.damping = o.damping;
  cfg.max_iterations = o.iters;
  cfg.rel_tol = o.relTol;
  cfg.abs_tol = o.absTol;
  cfg.residual_tol = o.residualTol;
  cfg.chebyshev_rho = o.chebRho;
  cfg.chebyshev_start = o.chebStart;
  int inst = 0;
  if ((rc = pdx_add_instance(ctx, &sc, &cfg, &inst))) return fail("pdx_add_instance", rc);

  pdx_diag d{};
  if ((rc = pdx_diagnostics(ctx, inst, &d))) return fail("pdx_diagnostics", rc);
  const double gnorm = std::sqrt(cfg.gravity[0] * cfg.gravity[0] + cfg.gravity[1] * cfg.gravity[1] +
                                 cfg.gravity[2] * cfg.gravity[2]);
  out.add("=== PD cloth benchmark (pdx, MI355X) ===\n");
  out.add("vertices %d  constraints %d  mass %.6g  device %d  precision %s  solve %s\n", s.n, s.E, d.total_mass,
          o.device, o.precision == PDX_PREC_FP64 ? "fp64" : "mixed32",
          o.solve == PDX_SOLVE_EXACT ? "exact" : (o.solve == PDX_SOLVE_FAST ? "fast" : "dense"));
  out.add("dt %.6g  substeps/frame %d  stiffness %.6g  PD iteration cap %d\n", o.dt, o.substeps, o.stiffness, o.iters);
  if (s.B || o.shear > 0.0)
    out.add("bend stencils %d (k %.6g)  shear stiffness %.6g\n", s.B, o.bend, o.shear);
  if (o.chebRho > 0.0)
    out.add("Chebyshev blend: rho %.6g, from iteration %d\n", o.chebRho, o.chebStart + 1);
  if (o.preset >= 0)
    out.add("iteration preset %s\n", kPresets[o.preset].name);
  else
    out.add("iteration preset (none: library default %d / %.3g)\n", 10, 1.0e-3);
  out.add("displacement tol %.3g%s   residual tol %.3g%s (unbalanced-force bound = tol x |g| = %.4g m/s^2)\n",
          o.relTol, o.residualTol > 0.0 ? " (not deciding)" : " (deciding)", o.residualTol,
          o.residualTol > 0.0 ? " (deciding)" : " (disabled)", o.residualTol > 0.0 ? o.residualTol * gnorm : 0.0);
  out.add("pin: %s\n", o.pinAll ? "every vertex (--pin-all)" : (s.fromMesh ? "none (loaded mesh)" : "two top corners (default)"));

  const auto f0 = std::chrono::steady_clock::now();
  if ((rc = pdx_factor(ctx))) return fail("pdx_factor", rc);
  const double setupSec = std::chrono::duration<double>(std::chrono::steady_clock::now() - f0).count();
  long long nnz = 0;
  pdx_matrix_nnz(ctx, inst, &nnz);
  out.add("(L has %d dof, nnz %lld; analysis, factorisation and upload took %.3f ms before the first substep)\n\n",
          3 * s.n, nnz, setupSec * 1e3);

  pdx_stats st{};
  double worstStrain = 0.0, deviceMs = 0.0;
  long long totalIters = 0;
  int maxUsed = 0, converged = 0;
  const auto w0 = std::chrono::steady_clock::now();
  for (int k = 0; k < o.steps; ++k) {
    if ((rc = pdx_step(ctx, 1))) return fail("pdx_step", rc);
    deviceMs += pdx_last_kernel_ms(ctx);
    if ((rc = pdx_get_stats(ctx, inst, &st))) return fail("pdx_get_stats", rc);
    totalIters += st.iterations_used;
    maxUsed = std::max(maxUsed, st.iterations_used);
    converged += st.converged != 0;
    if ((rc = pdx_diagnostics(ctx, inst, &d))) return fail("pdx_diagnostics", rc);
    worstStrain = std::max(worstStrain, d.max_strain);
    if (o.reportEvery > 0 && (k + 1) % o.reportEvery == 0)
      out.add("  step %6d  iterations %2d  strain(mean) %.4g  energy %.6g  max speed %.4g\n", k + 1,
              st.iterations_used, d.mean_strain, d.total_energy, d.max_speed);
  }
  const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - w0).count();
  const int steps = std::max(1, o.steps);

  out.add("--- results ---\n");
  out.add("substeps %d   total %.3f s   per substep %.3f ms   mean iterations %.2f   max iterations in one substep %d\n",
          o.steps, wall, 1000.0 * wall / steps, (double)totalIters / steps, maxUsed);
  {
    const bool resGate = o.residualTol > 0.0;
    const bool dispGate = o.absTol > 0.0 || o.relTol > 0.0;
    out.add("convergence: %d/%d substeps reached the criterion (criterion: %s); last residual %.6g m/s^2%s\n",
            converged, o.steps, resGate ? "residual" : (dispGate ? "displacement" : "none, iteration cap"),
            st.last_residual, resGate ? "" : " (residual criterion off: negative = not computed)");
  }
  out.add("final strain: max %.6g  mean %.6g\n", d.max_strain, d.mean_strain);
  out.add("run max strain %.6g\n", worstStrain);
  out.add("energy: elastic %.6g  total %.6g\n", d.elastic_energy, d.total_energy);

  out.add("\n--- stage times (device, ms, share of the substep kernel) ---\n");
  double phases = 0.0;
  for (int q = 0; q < 5; ++q) phases += st.phase_seconds[q];
  if (phases > 0.0) {
    const double tot = deviceMs * 1e-3;
    auto pct = [&](double v) { return tot > 0.0 ? 100.0 * v / tot : 0.0; };
    out.add("  predict+velocity   %10.3f  %5.1f%%\n", st.phase_seconds[3] * 1e3, pct(st.phase_seconds[3]));
    out.add("  project+scatter+residual gather %10.3f  %5.1f%%\n", st.phase_seconds[0] * 1e3, pct(st.phase_seconds[0]));
    if (st.phase_seconds[2] > 0.0)
      out.add("  global solve       %10.3f  %5.1f%%  (forward %.3f + backward %.3f)\n",
              (st.phase_seconds[1] + st.phase_seconds[2]) * 1e3, pct(st.phase_seconds[1] + st.phase_seconds[2]),
              st.phase_seconds[1] * 1e3, st.phase_seconds[2] * 1e3);
    else  // the grid and dense forms: one bucket, their sweeps overlap across workgroups
      out.add("  global solve       %10.3f  %5.1f%%  (both sweeps, overlapping across the grid)\n",
              st.phase_seconds[1] * 1e3, pct(st.phase_seconds[1]));
    out.add("  criterion          %10.3f  %5.1f%%\n", st.phase_seconds[4] * 1e3, pct(st.phase_seconds[4]));
  }
  out.add("  substep kernel     %10.3f  (%.4f ms per substep)\n", deviceMs, deviceMs / steps);
  out.add("  analysis+factor    %10.3f  (host, only when the stamp changes)\n",
          (st.analyze_seconds + st.factorize_seconds) * 1e3);

  out.add("\n--- factor reuse ---\n");
  out.add("  L dof %d  nnz %lld (structure)  factor nnz %lld\n", 3 * s.n, nnz, 3 * st.factor_nnz);
  out.add("  analyses %d  factorisations %d  solves %lld\n", st.analyze_count, st.factorize_count, st.total_iterations);
  out.add("  analysis time %.3f ms\n", st.analyze_seconds * 1e3);
  out.add("  factorisation %.3f ms\n", st.factorize_seconds * 1e3);
  out.add("  sweep steps per solve: forward %d  backward %d (levels %d / %d)\n", st.fwd_steps, st.bwd_steps,
          st.fwd_levels, st.bwd_levels);

  if (summary) {
    pdx_bench_summary& r = *summary;
    std::memset(&r, 0, sizeof(r));
    r.n_vertices = s.n, r.n_edges = s.E, r.n_bends = s.B, r.n_pinned = npin;
    r.substeps = o.steps, r.converged_substeps = converged, r.max_iterations_used = maxUsed;
    r.total_iterations = totalIters;
    r.last_residual = st.last_residual;
    r.max_strain = d.max_strain, r.mean_strain = d.mean_strain, r.run_max_strain = worstStrain;
    r.elastic_energy = d.elastic_energy, r.total_energy = d.total_energy;
    r.max_speed = d.max_speed, r.total_mass = d.total_mass;
    r.wall_seconds = wall;
    r.device_ms_per_substep = deviceMs / steps;
    r.analyze_count = st.analyze_count, r.factorize_count = st.factorize_count, r.color_count = st.color_count;
    r.matrix_nnz = nnz, r.factor_nnz = 3 * st.factor_nnz;
    r.analyze_seconds = st.analyze_seconds, r.factorize_seconds = st.factorize_seconds;
    r.loaded_mesh = s.fromMesh ? 1 : 0;
  }
  pdx_destroy(ctx);
  return finish(PDX_OK);
}
