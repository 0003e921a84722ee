// cavity_les.cpp — the drivers of the reference (bbrangeo/cfd_cavity
// src/main.rs:42-443 `run_uniform`, :445-738 `run_octree`, flags of
// src/cli.rs:19-384) on the GPU solvers of libcavlbm, through the C++ mirror
// in include/cavlbm.hpp.
//
// Same flags (--key val or --key=val), defaults, validation messages' intent,
// progress lines, abort guards and exit codes:
//   0 ok / help, 1 CLI error, 2 abort (max_u > 0.10 or non-finite),
//   3 no averaging samples, 4 output directory, 5 write failure,
//   6 snapshot read / restart failure.
// Outputs in --out: progress.log, snapshot.bin (CAVSNAP1; OCTSNAP1 with
// --mode octree); uniform also summary.json, fields_mean.csv, profiles_w.csv
// and snapshot.vti (post.rs:202-316).
// Extra flags: --storage fp64|fp32|fp16_offset, --geq-order 1|2, --device N.
#include <sys/stat.h>

#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cavlbm.hpp"

namespace {

struct Args {
  size_t nx = 64, ny = 256;
  double ra = 1.0e6, pr = 0.71, cs = 0.10, pr_t = 0.9;
  size_t steps = 200000, sample_start = 100000, sample_every = 100;
  std::string out = "output/runs_uniforme/ra1e6";
  double utarget = 0.02;
  uint64_t seed = 42;
  size_t progress_every = 1000, jobs = 1;
  double alpha_mult = 1.0, nu_start_mult = 1.0;
  std::string restart;
  double nu = -1.0;
  size_t snapshot_every = 0;
  bool dynamic_smag = false;
  double smag_alpha = 0.5, nu_sm_min = -0.1, eta_test = 2.0;
  bool octree = false;  // --mode octree
  size_t nx0 = 64, ny0 = 256, bl_width = 8, levels = 2;
  bool odd_exact = false;
  int storage = CAVLBM_STORAGE_FP64, geq_order = 2, device = 0;
};

constexpr double kMaxUAbort = 0.10;  // src/main.rs:46

void usage() {
  std::printf(
      "Usage: cavity_les [OPTIONS]   (uniform and octree modes of bbrangeo/cfd_cavity, MI355X)\n"
      "  --nx --ny --ra --pr --cs --pr_t --steps --sample-start --sample-every --out\n"
      "  --utarget --seed --progress-every --jobs --alpha-mult --nu-start-mult --restart\n"
      "  --nu --snapshot-every --dynamic-smag --smag-alpha --nu-sm-min --eta-test\n"
      "  --storage fp64|fp32|fp16_offset --geq-order 1|2 --device N\n"
      "  --mode uniform|octree; octree: --nx0 --ny0 --bl-width --levels 2|3 --odd-exact\n");
}

bool parse_size(const std::string& v, size_t& out) {
  char* e = nullptr;
  const unsigned long long x = std::strtoull(v.c_str(), &e, 10);
  if (v.empty() || *e || v[0] == '-') return false;
  out = (size_t)x;
  return true;
}
bool parse_f64(const std::string& v, double& out) {
  char* e = nullptr;
  out = std::strtod(v.c_str(), &e);
  return !v.empty() && !*e;
}

// src/cli.rs:190-384 (uniform subset), returns "" or an error message.
std::string parse(int argc, char** argv, Args& a, bool& help) {
  for (int i = 1; i < argc; ++i) {
    std::string arg = argv[i];
    if (arg == "-h" || arg == "--help") {
      help = true;
      return "";
    }
    std::string key = arg, val;
    bool inl = false;
    const size_t eq = arg.find('=');
    if (eq != std::string::npos) {
      key = arg.substr(0, eq);
      val = arg.substr(eq + 1);
      inl = true;
    }
    while (!key.empty() && key[0] == '-') key.erase(0, 1);
    for (auto& c : key)
      if (c == '_' && key != "pr_t") c = '-';
    if (key == "dynamic-smag") {
      a.dynamic_smag = !inl || val == "true" || val == "1";
      if (inl && val != "true" && val != "1" && val != "false" && val != "0")
        return "--dynamic-smag : booléen attendu, lu '" + val + "'";
      continue;
    }
    if (key == "odd-exact") {  // cli.rs:213-223
      a.odd_exact = !inl || val == "true" || val == "1";
      if (inl && val != "true" && val != "1" && val != "false" && val != "0")
        return "--odd-exact : booléen attendu, lu '" + val + "'";
      continue;
    }
    if (!inl) {
      if (i + 1 >= argc) return "valeur manquante pour --" + key;
      val = argv[++i];
    }
    bool ok = true;
    if (key == "mode") {
      ok = val == "uniform" || val == "octree";
      a.octree = val == "octree";
    } else if (key == "nx0") ok = parse_size(val, a.nx0);
    else if (key == "ny0") ok = parse_size(val, a.ny0);
    else if (key == "bl-width") ok = parse_size(val, a.bl_width);
    else if (key == "levels") ok = parse_size(val, a.levels);
    else if (key == "nx") ok = parse_size(val, a.nx);
    else if (key == "ny") ok = parse_size(val, a.ny);
    else if (key == "ra") ok = parse_f64(val, a.ra);
    else if (key == "pr") ok = parse_f64(val, a.pr);
    else if (key == "cs") ok = parse_f64(val, a.cs);
    else if (key == "pr_t" || key == "pr-t") ok = parse_f64(val, a.pr_t);
    else if (key == "steps") ok = parse_size(val, a.steps);
    else if (key == "sample-start") ok = parse_size(val, a.sample_start);
    else if (key == "sample-every") ok = parse_size(val, a.sample_every);
    else if (key == "out") a.out = val;
    else if (key == "utarget") ok = parse_f64(val, a.utarget);
    else if (key == "seed") {
      size_t s = 0;
      ok = parse_size(val, s);
      a.seed = s;
    } else if (key == "progress-every") ok = parse_size(val, a.progress_every);
    else if (key == "jobs") ok = parse_size(val, a.jobs);
    else if (key == "alpha-mult") ok = parse_f64(val, a.alpha_mult);
    else if (key == "nu-start-mult") ok = parse_f64(val, a.nu_start_mult);
    else if (key == "restart") a.restart = val;
    else if (key == "snapshot-every") ok = parse_size(val, a.snapshot_every);
    else if (key == "nu") {
      ok = parse_f64(val, a.nu);
      if (ok && !(a.nu > 0.0 && std::isfinite(a.nu))) return "--nu doit être > 0 et fini";
      if (ok && a.nu >= 0.002)
        return "--nu " + val + " ≥ 0.002 : inutile, le calage utarget suffit ; --nu est réservé au sous-plancher";
    } else if (key == "storage") {
      if (val == "fp64") a.storage = CAVLBM_STORAGE_FP64;
      else if (val == "fp32") a.storage = CAVLBM_STORAGE_FP32;
      else if (val == "fp16_offset") a.storage = CAVLBM_STORAGE_FP16_OFFSET;
      else ok = false;
    } else if (key == "geq-order") {
      size_t g = 0;
      ok = parse_size(val, g) && (g == 1 || g == 2);
      a.geq_order = (int)g;
    } else if (key == "device") {
      size_t d = 0;
      ok = parse_size(val, d);
      a.device = (int)d;
    } else if (key == "smag-alpha") ok = parse_f64(val, a.smag_alpha);
    else if (key == "nu-sm-min") ok = parse_f64(val, a.nu_sm_min);
    else if (key == "eta-test") ok = parse_f64(val, a.eta_test);
    else {
      return "option inconnue : --" + key;
    }
    if (!ok) return "--" + key + " : valeur invalide '" + val + "'";
  }
  // src/cli.rs:334-356
  if (a.nx < 8 || a.ny < 8) return "nx et ny doivent être ≥ 8";
  if (a.ny <= a.nx) return "cavité haute : ny doit être > nx (Az ≈ 4)";
  if (a.ra <= 0.0 || a.pr <= 0.0) return "ra et pr doivent être > 0";
  if (a.sample_every == 0) return "sample-every doit être ≥ 1";
  if (a.jobs == 0) return "jobs doit être ≥ 1";
  if (!(a.alpha_mult > 0.0) || !std::isfinite(a.alpha_mult)) return "alpha-mult doit être > 0 et fini";
  if (!(a.nu_start_mult >= 1.0) || !std::isfinite(a.nu_start_mult))
    return "nu-start-mult doit être ≥ 1 et fini";
  if (!(a.smag_alpha >= 0.0 && a.smag_alpha <= 1.0) || !std::isfinite(a.smag_alpha))
    return "smag-alpha doit être ∈ [0, 1] (0 = Smagorinsky pur, 1 = similarité pure)";
  if (!(a.eta_test >= 1.0) || !std::isfinite(a.eta_test))
    return "eta-test doit être ≥ 1 (ratio filtre test/grille)";
  if (a.octree) {  // cli.rs:362-381
    if (a.nx0 < 16 || a.ny0 < 16) return "nx0 et ny0 doivent être ≥ 16";
    if (a.ny0 <= a.nx0) return "cavité haute : ny0 doit être > nx0 (Az ≈ 4)";
    if (!(a.bl_width >= 2 && 2 * a.bl_width < a.nx0))
      return "bl-width doit être ≥ 2 et 2·bl-width < nx0 (noyau grossier requis)";
    if (!(a.levels == 2 || a.levels == 3))
      return "levels doit valoir 2 (config V2) ou 3 (escalier tanh V6)";
    if (a.levels == 3 && a.bl_width != 4)
      std::fprintf(stderr,
                   "Note : escalier V6 conçu pour bl-width = 4 cellules grossières (frontières "
                   "tanh §0.1) ; lu %zu\n",
                   a.bl_width);
  }
  if (a.progress_every == 0) a.progress_every = 1;
  return "";
}

// Rust's {:.3e}: mantissa, 'e', exponent without '+' or zero padding.
std::string rust_exp(double v, int prec) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.*e", prec, v);
  std::string s(buf);
  const size_t e = s.find('e');
  if (e == std::string::npos) return s;
  const int ex = std::atoi(s.c_str() + e + 1);
  return s.substr(0, e) + "e" + std::to_string(ex);
}

// Rust's {} for f64: the shortest decimal that reads back to the same value,
// positional (no exponent), without a trailing ".0".
std::string rust_display(double v) {
  if (!std::isfinite(v)) return std::isnan(v) ? "NaN" : (v > 0 ? "inf" : "-inf");
  if (v == 0.0) return std::signbit(v) ? "-0" : "0";
  char buf[64];
  int prec = 1;
  for (; prec <= 17; ++prec) {
    std::snprintf(buf, sizeof buf, "%.*e", prec - 1, v);
    if (std::strtod(buf, nullptr) == v) break;
  }
  std::string m(buf);
  const size_t e = m.find('e');
  const int ex = std::atoi(m.c_str() + e + 1);
  std::string digits;
  for (size_t i = 0; i < e; ++i)
    if (std::isdigit((unsigned char)m[i])) digits += m[i];
  while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
  std::string out = v < 0 ? "-" : "";
  const int point = ex + 1;  // digits before the decimal point
  if (point <= 0) {
    out += "0." + std::string(-point, '0') + digits;
  } else if ((size_t)point >= digits.size()) {
    out += digits + std::string(point - digits.size(), '0');
  } else {
    out += digits.substr(0, point) + "." + digits.substr(point);
  }
  return out;
}

// main.rs:740-753
// stream_function_mean (main.rs:723-737): trapezoidal integration of the mean
// u_x up each column, psi = 0 on the bottom row.
std::vector<double> stream_function_mean(const std::vector<double>& ux, size_t nx, size_t ny) {
  std::vector<double> psi(nx * ny, 0.0);
  for (size_t ix = 0; ix < nx; ++ix) {
    double acc = 0.0;
    for (size_t iy = 1; iy < ny; ++iy) {
      acc += 0.5 * (ux[(iy - 1) * nx + ix] + ux[iy * nx + ix]);
      psi[iy * nx + ix] = acc;
    }
  }
  return psi;
}

// post::write_fields_csv (post.rs:202-232): x, y, ux, uy, T, psi, nu_t of the
// mean fields, row by row.
std::string fields_csv(size_t nx, size_t ny, const std::vector<double>& ux,
                       const std::vector<double>& uy, const std::vector<double>& t,
                       const std::vector<double>& psi, const std::vector<double>& nut) {
  std::string out = "x,y,ux,uy,T,psi,nu_t\n";
  out.reserve(out.size() + nx * ny * 96);
  char b[64];
  for (size_t iy = 0; iy < ny; ++iy)
    for (size_t ix = 0; ix < nx; ++ix) {
      const size_t k = iy * nx + ix;
      std::snprintf(b, sizeof b, "%.6f,%.6f", (double)ix, (double)iy);
      out += b;
      for (const std::vector<double>* v : {&ux, &uy, &t, &psi, &nut}) {
        out += ',';
        out += rust_exp((*v)[k], 8);
      }
      out += '\n';
    }
  return out;
}

// post::write_profiles_csv (post.rs:235-249): w at z = 0.5 and 0.75 of the
// height along x / (nx - 1).
std::string profiles_csv(const std::vector<double>& mid, const std::vector<double>& q75,
                         size_t nx) {
  std::string out = "x_norm,w_z05,w_z075\n";
  char b[32];
  for (size_t ix = 0; ix < nx; ++ix) {
    std::snprintf(b, sizeof b, "%.6f", (double)ix / (double)(nx - 1));
    out += b;
    out += ',' + rust_exp(mid[ix], 8) + ',' + rust_exp(q75[ix], 8) + '\n';
  }
  return out;
}

// post::write_vtk (post.rs:252-316): VTK XML ImageData of the final T, ux, uy,
// x varying fastest (k = iy*nx + ix), one line per row.
std::string vtk_image(size_t nx, size_t ny, const std::vector<double>& t,
                      const std::vector<double>& ux, const std::vector<double>& uy) {
  std::string out = "<?xml version=\"1.0\"?>\n"
                    "<VTKFile type=\"ImageData\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
  char b[160];
  std::snprintf(b, sizeof b,
                "  <ImageData WholeExtent=\"0 %zu 0 %zu 0 0\" Origin=\"0 0 0\" Spacing=\"1 1 1\">\n"
                "    <Piece Extent=\"0 %zu 0 %zu 0 0\">\n",
                nx - 1, ny - 1, nx - 1, ny - 1);
  out += b;
  auto block = [&](const std::vector<double>& v) {
    std::string s;
    s.reserve(nx * ny * 14);
    for (size_t iy = 0; iy < ny; ++iy) {
      for (size_t ix = 0; ix < nx; ++ix) s += rust_exp(v[iy * nx + ix], 6) + ' ';
      s += '\n';
    }
    return s;
  };
  out += "      <PointData Scalars=\"Temperature\">\n";
  out += "        <DataArray type=\"Float64\" Name=\"Temperature\" format=\"ascii\">\n";
  out += block(t) + "        </DataArray>\n";
  out += "        <DataArray type=\"Float64\" Name=\"Ux\" format=\"ascii\">\n";
  out += block(ux) + "        </DataArray>\n";
  out += "        <DataArray type=\"Float64\" Name=\"Uy\" format=\"ascii\">\n";
  out += block(uy) + "        </DataArray>\n";
  out += "      </PointData>\n      <CellData>\n      </CellData>\n    </Piece>\n"
         "  </ImageData>\n</VTKFile>\n";
  return out;
}

double nusselt_from_mean(const std::vector<double>& t, size_t nx, size_t ny, double lz, bool hot) {
  double sum = 0.0;
  for (size_t iy = 0; iy < ny; ++iy) {
    const double g = hot ? t[iy * nx + 1] - t[iy * nx + 2] : t[iy * nx + nx - 3] - t[iy * nx + nx - 2];
    sum += std::fabs(g);
  }
  return sum / (double)ny * lz / 1.0;
}

// main.rs:756-770 (max_of)
std::pair<double, double> max_of(const std::vector<double>& prof, size_t nx) {
  double best = 0.0;
  size_t bi = 1;
  for (size_t i = 1; i + 1 < nx; ++i)
    if (std::fabs(prof[i]) > best) {
      best = std::fabs(prof[i]);
      bi = i;
    }
  return {best, (double)bi / (double)(nx - 1)};
}

bool write_text(const std::string& path, const std::string& text) {
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) return false;
  const size_t w = std::fwrite(text.data(), 1, text.size(), f);
  return std::fclose(f) == 0 && w == text.size();
}

bool make_dirs(const std::string& path) {
  std::string cur;
  for (size_t i = 0; i <= path.size(); ++i) {
    if (i == path.size() || path[i] == '/') {
      if (!cur.empty()) {
        struct stat st;
        if (stat(cur.c_str(), &st) != 0 && mkdir(cur.c_str(), 0755) != 0) return false;
      }
    }
    if (i < path.size()) cur.push_back(path[i]);
  }
  return true;
}

}  // namespace

std::string fixed(double v, int prec) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.*f", prec, v);
  return buf;
}

// run_octree (main.rs:445-738): progress lines with the interface noise and
// the Nusselt estimators, abort guards, sampled Nusselt means, OCTSNAP1
// exports and restart, and the V2 verdict against the uniform references.
int run_octree(const Args& a) {
  std::printf("=== Cavité DHC 2D octree (Étape B, porte V2) ===\n");
  std::printf("Grossier : %zu×%zu | BL fine : %zu cellules grossières ×2 parois (équiv. uniforme "
              "%zu×%zu) | Ra = %s | Pr = %s | jobs = %zu | ξ_odd = %s\n",
              a.nx0, a.ny0, a.bl_width, 2 * a.bl_width, a.ny0, rust_exp(a.ra, 3).c_str(),
              rust_display(a.pr).c_str(), a.jobs, a.odd_exact ? "exact (τ⁻)" : "1/2 (approx.)");
  if (a.nu > 0.0)
    std::printf("Étape H : nu forcé = %s (bypass calage utarget = %s)\n",
                rust_display(a.nu).c_str(), rust_display(a.utarget).c_str());
  else if (a.utarget != 0.02)
    std::printf("Étape H : calage ν via utarget = %s\n", rust_display(a.utarget).c_str());
  if (a.nu_start_mult > 1.0) {
    std::printf("Étape H : rampe de viscosité exponentielle nu-start-mult = %s (tau_ramp = 1000 "
                "pas, terminée ~5000 pas) + rampe beta linéaire sur %zu pas (spin-up)\n",
                rust_display(a.nu_start_mult).c_str(), std::max<size_t>(a.ny0 / 10, 2000));
    if (a.sample_start < 5000)
      std::fprintf(stderr,
                   "WARNING : rampe de viscosité terminée ~5000 pas mais sample-start = %zu : "
                   "augmentez --sample-start (> 5000) pour ne pas moyenner pendant la rampe\n",
                   a.sample_start);
  }
  try {
    cavlbm::OctreeCavitySolver sim(a.nx0, a.ny0, a.bl_width, a.ra, a.pr, a.utarget, a.seed,
                                   a.jobs, a.odd_exact, (int)a.levels, a.nu_start_mult, a.nu,
                                   a.geq_order, a.device);
    // main.rs:515-523
    sim.set_dynamic_sgs(a.dynamic_smag, a.smag_alpha, a.nu_sm_min, a.eta_test);
    if (a.dynamic_smag)
      std::printf("LES dynamique (Étape D) : alpha_mixte = %s | nu_sm_min = %s | eta_test = %s\n",
                  rust_display(a.smag_alpha).c_str(), rust_display(a.nu_sm_min).c_str(),
                  rust_display(a.eta_test).c_str());
    const int ng = sim.n_grids();
    if (!a.restart.empty()) {  // main.rs:526-546
      uint32_t lv = 0, sng = 0;
      uint64_t step = 0, dims[10];
      if (cavlbm_octsnap_read(a.restart.c_str(), &lv, &step, &sng, dims, nullptr, nullptr, nullptr,
                              nullptr, 0)) {
        std::fprintf(stderr, "Erreur de lecture du snapshot %s : %s\n", a.restart.c_str(),
                     cavlbm_snapshot_last_error());
        return 6;
      }
      uint64_t total = 0;
      for (uint32_t g = 0; g < sng; ++g) total += dims[2 * g] * dims[2 * g + 1];
      std::vector<double> f[4];
      for (auto& v : f) v.resize(total);
      if (cavlbm_octsnap_read(a.restart.c_str(), &lv, &step, &sng, dims, f[0].data(), f[1].data(),
                              f[2].data(), f[3].data(), total)) {
        std::fprintf(stderr, "Erreur de lecture du snapshot %s : %s\n", a.restart.c_str(),
                     cavlbm_snapshot_last_error());
        return 6;
      }
      std::printf("Restart depuis %s : snapshot levels=%u (%u grilles), pas %llu — injection des "
                  "champs du solveur octree\n",
                  a.restart.c_str(), lv, sng, (unsigned long long)step);
      // set_from_snapshot's layout checks (solver.rs:1095-1133)
      if ((int)lv != (int)a.levels || (int)sng != ng) {
        std::fprintf(stderr, "Erreur de restart : restart : snapshot levels=%u (%u grilles) "
                     "incompatible avec le solveur levels=%zu (%d grilles attendues)\n",
                     lv, sng, a.levels, ng);
        return 6;
      }
      for (int g = 0; g < ng; ++g)
        if (dims[2 * g] != sim.dims()[g].first || dims[2 * g + 1] != sim.dims()[g].second) {
          std::fprintf(stderr, "Erreur de restart : restart : grille %d du snapshot %llux%llu ≠ "
                       "solveur %zux%zu (même layout requis — pas d'interpolation inter-grille "
                       "en Étape G)\n",
                       g, (unsigned long long)dims[2 * g], (unsigned long long)dims[2 * g + 1],
                       sim.dims()[g].first, sim.dims()[g].second);
          return 6;
        }
      sim.set_from_snapshot(f[0], f[1], f[2], f[3]);
    }
    const cavlbm_octree_derived& d = sim.derived();
    const size_t n0 = a.nx0 * a.ny0;
    std::printf("Unités réseau : nu = %.6f | alpha = %.6f | tau+ = %.4f/%.4f%s | beta_eff = "
                "%s/%s%s | %zu niveaux | %llu nœuds (grossier %zu + fins %llu)\n",
                d.nu_lu, d.alpha_lu, d.tau[0][0], d.tau[1][0],
                a.levels == 3 ? ("/" + fixed(d.tau[2][0], 4)).c_str() : "",
                rust_exp(d.beta_eff_coarse, 4).c_str(), rust_exp(d.beta_eff_fine, 4).c_str(),
                a.levels == 3 ? ("/" + rust_exp(0.25 * d.beta_eff_coarse, 4)).c_str() : "",
                a.levels, (unsigned long long)d.n_nodes_total, n0,
                (unsigned long long)(d.n_nodes_total - n0));
    const double lz = (double)(a.ny0 - 1);
    std::printf("Ra vérifié depuis unités réseau (grossier) : %s (cible %s)\n",
                rust_exp(d.beta_eff_coarse * (lz * lz * lz) / (d.nu_lu * d.alpha_lu), 4).c_str(),
                rust_exp(a.ra, 4).c_str());
    if (!make_dirs(a.out)) {
      std::fprintf(stderr, "Impossible de créer %s\n", a.out.c_str());
      return 4;
    }
    std::vector<std::string> log_lines;
    double nuc_sum = 0.0, nuf_sum = 0.0, nucl0_sum = 0.0, nufine_sum = 0.0;
    size_t nuc_n = 0;
    double noise_max = 0.0, noise_iface_max = 0.0;
    const std::string snap_path = a.out + "/snapshot.bin";
    auto write_snap = [&](uint64_t step) {
      std::vector<double> r, u, v, t;
      sim.level_fields(r, u, v, t);
      std::vector<uint64_t> dims;
      for (auto& p : sim.dims()) {
        dims.push_back(p.first);
        dims.push_back(p.second);
      }
      return cavlbm_octsnap_write(snap_path.c_str(), (uint32_t)a.levels, step, (uint32_t)ng,
                                  dims.data(), r.data(), u.data(), v.data(), t.data(), 1);
    };
    auto write_log = [&]() {
      std::string text;
      for (size_t i = 0; i < log_lines.size(); ++i) text += log_lines[i] + "\n";
      write_text(a.out + "/progress.log", text);
    };
    const auto t0 = std::chrono::steady_clock::now();
    size_t done = 0;
    auto is_progress = [&](size_t s) { return s % a.progress_every == 0 || s + 1 == a.steps; };
    auto is_sample = [&](size_t s) {
      return s >= a.sample_start && (s - a.sample_start) % a.sample_every == 0;
    };
    auto is_snap = [&](size_t s) { return a.snapshot_every > 0 && (s + 1) % a.snapshot_every == 0; };
    for (size_t step = 0; step < a.steps; ++step) {
      if (!(is_progress(step) || is_sample(step) || is_snap(step))) continue;
      sim.run(step + 1 - done);
      done = step + 1;
      cavlbm_octree_diag dg{};
      if (is_sample(step) || is_progress(step)) dg = sim.diagnostics();
      if (is_sample(step)) {
        nuc_sum += dg.nusselt_hot;
        nuf_sum += dg.nusselt_cold;
        nucl0_sum += dg.nusselt_hot_link0;
        nufine_sum += dg.nusselt_hot_fine;
        ++nuc_n;
      }
      if (is_progress(step)) {
        noise_max = std::fmax(noise_max, dg.max_rho_dev);
        noise_iface_max = std::fmax(noise_iface_max, dg.max_rho_dev_ifaces);
        // main.rs:627-641: the model's diagnostics as a suffix
        std::string dyn_suffix;
        if (a.dynamic_smag) {
          const std::vector<double> st = sim.dynamic_stats();
          char b[160];
          std::snprintf(b, sizeof b, " | cs_dyn = %.4f | kap_dyn = %.4f | nu_t_max = %s", st[0],
                        st[2], rust_exp(st[4], 3).c_str());
          dyn_suffix = b;
        }
        char line[640];
        std::snprintf(line, sizeof line,
                      "step %7zu | max_rho_dev = %s | max_rho_dev_ifaces = %s | masse = %.6f | "
                      "Nuc = %.4f | Nuc_link0 = %.4f | Nuf = %.4f | Nu_fin = %.4f | max_u = %.5f%s",
                      step, rust_exp(dg.max_rho_dev, 3).c_str(),
                      rust_exp(dg.max_rho_dev_ifaces, 3).c_str(), dg.masse, dg.nusselt_hot,
                      dg.nusselt_hot_link0, dg.nusselt_cold, dg.nusselt_hot_fine, dg.max_u_all,
                      dyn_suffix.c_str());
        std::printf("%s\n", line);
        log_lines.push_back(line);
        if (dg.guard != 0) {  // check_guards (solver.rs:1018-1041)
          char m[256];
          if (dg.guard == 1)
            std::snprintf(m, sizeof m,
                          "AVORT : max_u = %.5f dépasse le garde-fou 0.10 (Ma_local = max_u·√3 "
                          "≈ %.3f)",
                          dg.max_u_all, dg.max_u_all * std::sqrt(3.0));
          else
            std::snprintf(m, sizeof m, "DIVERGENCE détectée (valeurs non finies).");
          std::fprintf(stderr, "%s\n", m);
          log_lines.push_back(m);
          write_log();
          return 2;
        }
      }
      if (is_snap(step)) {
        if (write_snap(step)) {
          std::fprintf(stderr, "Erreur d'export du snapshot au pas %zu : %s\n", step,
                       cavlbm_snapshot_last_error());
          return 5;
        }
        std::printf("snapshot exporté au pas %zu\n", step);
      }
    }
    if (done < a.steps) sim.run(a.steps - done);
    if (a.snapshot_every > 0) {
      if (write_snap(a.steps)) {
        std::fprintf(stderr, "Erreur d'export du snapshot final : %s\n",
                     cavlbm_snapshot_last_error());
        return 5;
      }
      std::printf("snapshot final exporté au pas %zu\n", a.steps - 1);
    }
    const double el =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("Simulation terminée en %.1f s (%.1fk pas/s)\n", el, (double)a.steps / el / 1000.0);
    if (nuc_n == 0) {
      std::fprintf(stderr, "Aucun échantillon moyenné : sample_start ≥ steps ?\n");
      return 3;
    }
    const double nuc_mean = nuc_sum / (double)nuc_n, nuf_mean = nuf_sum / (double)nuc_n;
    const double nucl0_mean = nucl0_sum / (double)nuc_n, nufine_mean = nufine_sum / (double)nuc_n;
    const double nuc_ref = 2.3585, nucl0_ref = 1.3660;  // main.rs:697-700
    const double ecart = std::fabs((nuc_mean - nuc_ref) / nuc_ref) * 100.0;
    char summ[1024];
    std::snprintf(summ, sizeof summ,
                  "--- Bilan (moyenné sur %zu échantillons, fenêtre %zu→%zu) ---\n"
                  "Nuc (2ᵉ liaison) = %.4f | écart like-for-like vs %s = %.1f %%\n"
                  "Nuc_link0 (1ʳᵉ liaison) = %.4f | Nuf = %.4f | Nu_fin (bloc chaud) = %.4f\n"
                  "Références uniforme Ra=1e6 : 2ᵉ liaison = %s | 1ʳᵉ liaison = %s\n"
                  "Bruit acoustique (max sur la run) : max|ρ−ρ̄| = %s | interfaces = %s\n"
                  "PORTE V2 (écart Nuc ≤ 5 %%) : %s",
                  nuc_n, a.sample_start, a.steps, nuc_mean, rust_display(nuc_ref).c_str(), ecart,
                  nucl0_mean, nuf_mean, nufine_mean, rust_display(nuc_ref).c_str(),
                  rust_display(nucl0_ref).c_str(), rust_exp(noise_max, 3).c_str(),
                  rust_exp(noise_iface_max, 3).c_str(), ecart <= 5.0 ? "PASS" : "FAIL");
    std::printf("%s\n", summ);
    log_lines.push_back(summ);
    write_log();
    std::printf("Artefacts écrits dans %s\n", a.out.c_str());
  } catch (const std::invalid_argument& e) {
    std::fprintf(stderr, "Erreur : %s\n", e.what());
    return 1;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "Erreur : %s\n", e.what());
    return 1;
  }
  return 0;
}

int main(int argc, char** argv) {
  Args a;
  bool help = false;
  const std::string err = parse(argc, argv, a, help);
  if (help) {
    usage();
    return 0;
  }
  if (!err.empty()) {
    std::fprintf(stderr, "Erreur : %s\n", err.c_str());
    usage();
    return 1;
  }
  if (a.octree) return run_octree(a);  // main.rs:34-37
  std::printf("=== Cavité différentiellement chauffée 2D (LBM D2Q9 + Smagorinsky, MI355X) ===\n");
  std::printf("Grille : %zux%zu (Az = %.2f) | Ra = %.3e | Pr = %g | Cs = %g | Pr_t = %g | jobs = %zu\n",
              a.nx, a.ny, (double)(a.ny - 1) / (double)(a.nx - 1), a.ra, a.pr, a.cs, a.pr_t, a.jobs);
  try {
    cavlbm::CavitySolver sim(a.nx, a.ny, a.ra, a.pr, a.cs, a.pr_t, a.utarget, a.seed, a.jobs,
                             a.alpha_mult, a.nu_start_mult, a.nu, a.storage, a.geq_order,
                             a.device);
    // main.rs:94-102
    sim.set_dynamic_sgs(a.dynamic_smag, a.smag_alpha, a.nu_sm_min, a.eta_test);
    if (a.dynamic_smag)
      std::printf("LES dynamique : échelles mixtes α = %s (Tab. 2.3) | η_test = %s | νsm_min = %s·ν\n",
                  rust_display(a.smag_alpha).c_str(), rust_display(a.eta_test).c_str(),
                  rust_display(a.nu_sm_min).c_str());
    if (!a.restart.empty()) {  // main.rs:104-128
      uint64_t snx = 0, sny = 0;
      if (cavlbm_snapshot_read(a.restart.c_str(), &snx, &sny, nullptr, nullptr, nullptr, nullptr, 0)) {
        std::fprintf(stderr, "Erreur de lecture du snapshot %s : %s\n", a.restart.c_str(),
                     cavlbm_snapshot_last_error());
        return 6;
      }
      const size_t ns = snx * sny, n = a.nx * a.ny;
      std::vector<double> f[4], g[4];
      for (int q = 0; q < 4; ++q) {
        f[q].resize(ns);
        g[q].resize(n);
      }
      if (cavlbm_snapshot_read(a.restart.c_str(), &snx, &sny, f[0].data(), f[1].data(),
                               f[2].data(), f[3].data(), ns)) {
        std::fprintf(stderr, "Erreur de lecture du snapshot %s : %s\n", a.restart.c_str(),
                     cavlbm_snapshot_last_error());
        return 6;
      }
      std::printf("Restart depuis %s : snapshot %llux%llu → grille %zux%zu%s\n", a.restart.c_str(),
                  (unsigned long long)snx, (unsigned long long)sny, a.nx, a.ny,
                  (snx == a.nx && sny == a.ny) ? "" : " (interpolation bilinéaire)");
      for (int q = 0; q < 4; ++q)
        if (cavlbm_interpolate_field(f[q].data(), snx, sny, g[q].data(), a.nx, a.ny)) {
          std::fprintf(stderr, "Erreur de lecture du snapshot %s : %s\n", a.restart.c_str(),
                       cavlbm_snapshot_last_error());
          return 6;
        }
      sim.init_from_fields(g[0], g[1], g[2], g[3]);
    }
    const cavlbm_derived d = sim.derived();
    std::printf("Unités réseau : nu = %.6f | alpha = %.6f | tau = %.4f | tau_t = %.4f | beta_eff = %s\n",
                d.nu, d.alpha, d.tau, d.tau_t, rust_exp(d.beta_eff, 4).c_str());

    const size_t n_nodes = a.nx * a.ny;
    std::vector<std::string> log_lines;
    const std::string snap_path = a.out + "/snapshot.bin";
    if (a.snapshot_every > 0 && !make_dirs(a.out)) {
      std::fprintf(stderr, "Impossible de créer %s\n", a.out.c_str());
      return 4;
    }
    const auto t0 = std::chrono::steady_clock::now();
    size_t done = 0;
    auto is_progress = [&](size_t s) { return s % a.progress_every == 0 || s + 1 == a.steps; };
    auto is_sample = [&](size_t s) {
      return s >= a.sample_start && (s - a.sample_start) % a.sample_every == 0;
    };
    auto is_snap = [&](size_t s) { return a.snapshot_every > 0 && (s + 1) % a.snapshot_every == 0; };
    for (size_t step = 0; step < a.steps; ++step) {
      if (!(is_progress(step) || is_sample(step) || is_snap(step))) continue;
      sim.run(step + 1 - done);  // steps done..step issued back to back on the device
      done = step + 1;
      if (is_progress(step)) {  // main.rs:168-276
        const cavlbm_diag dg = sim.diagnostics();
        char line[384];
        char dyn[160] = "";
        if (a.dynamic_smag) {  // main.rs:179-187
          double st[4];
          cavlbm_dynamic_stats(sim.raw(), st);
          std::snprintf(dyn, sizeof dyn,
                        " | cs_dyn_mean = %.4f | cs_dyn_max = %.4f | kap_dyn_mean = %.4f", st[0],
                        st[1], st[2]);
        }
        std::snprintf(line, sizeof line,
                      "step %7zu | w_max = %.5f | max_u = %.5f | Nuc = %.4f | Nuf = %.4f | "
                      "masse = %.6f | nu_t_max = %s%s",
                      step, dg.max_uy, dg.max_u, dg.nusselt_hot, dg.nusselt_cold,
                      dg.total_mass / (double)n_nodes, rust_exp(dg.nu_t_max, 3).c_str(), dyn);
        std::printf("%s\n", line);
        log_lines.push_back(line);
        if (dg.max_u > kMaxUAbort) {
          char m[256];
          std::snprintf(m, sizeof m,
                        "AVORT : max_u = %.5f dépasse le garde-fou %.2f (Ma_local = max_u·√3 ≈ "
                        "%.3f > marge de sécurité)",
                        dg.max_u, kMaxUAbort, dg.max_u * std::sqrt(3.0));
          std::fprintf(stderr, "%s\n", m);
          log_lines.push_back(m);
          make_dirs(a.out);
          std::string text;
          for (auto& l : log_lines) text += l + "\n";
          write_text(a.out + "/progress.log", text);
          return 2;
        }
        if (!dg.all_finite) {
          std::fprintf(stderr, "DIVERGENCE détectée au pas %zu (valeurs non finies).\n", step);
          return 2;
        }
      }
      if (is_sample(step)) sim.average_accumulate();
      if (is_snap(step)) {  // main.rs:284-300
        const auto rho = sim.rho(), ux = sim.ux(), uy = sim.uy(), t = sim.temperature();
        if (cavlbm_snapshot_write_atomic(snap_path.c_str(), a.nx, a.ny, rho.data(), ux.data(),
                                         uy.data(), t.data())) {
          std::fprintf(stderr, "Erreur d'export du snapshot au pas %zu : %s\n", step,
                       cavlbm_snapshot_last_error());
          return 5;
        }
        std::printf("snapshot exporté au pas %zu\n", step);
      }
    }
    if (done < a.steps) sim.run(a.steps - done);
    const double el =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("Simulation terminée en %.1f s (%.1fk pas/s)\n", el, (double)a.steps / el / 1000.0);
    if (sim.average_count() == 0) {
      std::fprintf(stderr, "Aucun échantillon moyenné : sample_start ≥ steps ?\n");
      return 3;
    }
    std::printf("Moyennage sur %llu échantillons.\n", (unsigned long long)sim.average_count());
    std::vector<double> ux_m, uy_m, t_m, nut_m, uu_m;
    sim.average_means(ux_m, uy_m, t_m, nut_m, uu_m);
    const double nuc = nusselt_from_mean(t_m, a.nx, a.ny, d.lz, true);
    const double nuf = nusselt_from_mean(t_m, a.nx, a.ny, d.lz, false);
    std::vector<double> mid(a.nx), q75(a.nx);
    for (size_t ix = 0; ix < a.nx; ++ix) {
      mid[ix] = uy_m[(a.ny / 2) * a.nx + ix];
      q75[ix] = uy_m[(3 * a.ny / 4) * a.nx + ix];
    }
    const auto wm = max_of(mid, a.nx), w75 = max_of(q75, a.nx);
    const cavlbm_diag dg = sim.diagnostics();
    double s2 = 0.0;
    for (size_t k = 0; k < n_nodes; ++k)
      s2 += std::fmax(uu_m[k] - ux_m[k] * ux_m[k] - uy_m[k] * uy_m[k], 0.0);
    const double s_half = std::sqrt(s2 / (double)n_nodes);
    if (!make_dirs(a.out)) {
      std::fprintf(stderr, "Impossible de créer %s\n", a.out.c_str());
      return 4;
    }
    const auto rho = sim.rho(), ux = sim.ux(), uy = sim.uy(), t = sim.temperature();
    char js[2048];
    std::snprintf(js, sizeof js,
                  "{\n  \"case\": \"differentially_heated_cavity_2d\",\n  \"ra\": %s,\n"
                  "  \"cs\": %g,\n  \"grid\": [{\"nx\": %zu, \"ny\": %zu}],\n"
                  "  \"aspect_ratio\": %.4f,\n  \"steps\": %zu,\n  \"time_samples\": %llu,\n"
                  "  \"nusselt_hot\": %.4f,\n  \"nusselt_cold\": %.4f,\n  \"nusselt_volume\": %.4f,\n"
                  "  \"wmax_z05\": %.5f,\n  \"x_wmax_z05\": %.5f,\n  \"wmax_z075\": %.5f,\n"
                  "  \"x_wmax_z075\": %.5f,\n  \"y_plus_hot\": %.3f,\n  \"s_half\": %.4f\n}\n",
                  rust_exp(a.ra, 6).c_str(), a.cs, a.nx, a.ny,
                  (double)(a.ny - 1) / (double)(a.nx - 1), a.steps,
                  (unsigned long long)sim.average_count(), nuc, nuf, dg.nusselt_volume, wm.first,
                  wm.second, w75.first, w75.second, dg.y_plus_hot, s_half);
    std::string text;
    for (auto& l : log_lines) text += l + "\n";
    // main.rs:317-392: fields_mean.csv, profiles_w.csv, snapshot.vti,
    // snapshot.bin, summary.json, progress.log
    const std::vector<double> psi_m = stream_function_mean(ux_m, a.nx, a.ny);
    if (!write_text(a.out + "/fields_mean.csv",
                    fields_csv(a.nx, a.ny, ux_m, uy_m, t_m, psi_m, nut_m)) ||
        !write_text(a.out + "/profiles_w.csv", profiles_csv(mid, q75, a.nx)) ||
        !write_text(a.out + "/snapshot.vti", vtk_image(a.nx, a.ny, t, ux, uy)) ||
        cavlbm_snapshot_write(snap_path.c_str(), a.nx, a.ny, rho.data(), ux.data(), uy.data(),
                              t.data()) ||
        !write_text(a.out + "/summary.json", js) || !write_text(a.out + "/progress.log", text)) {
      std::fprintf(stderr, "Erreur d'écriture dans %s\n", a.out.c_str());
      return 5;
    }
    if (a.dynamic_smag) {  // main.rs:393-424, a companion file beside summary.json
      double st[4];
      cavlbm_dynamic_stats(sim.raw(), st);
      double nsum = 0.0, nmax = 0.0;
      for (double v : nut_m) {
        nsum += v;
        nmax = std::fmax(nmax, std::fabs(v));
      }
      char dj[1024];
      std::snprintf(dj, sizeof dj,
                    "{\n  \"dynamic_smag\": true,\n  \"smag_alpha\": %s,\n  \"eta_test\": %s,\n"
                    "  \"nu_sm_min\": %s,\n  \"cs_dyn_mean\": %.6f,\n  \"cs_dyn_max\": %.6f,\n"
                    "  \"kappa_dyn_mean\": %.6f,\n  \"kappa_dyn_max\": %.6f,\n"
                    "  \"nu_t_mean\": %s,\n  \"nu_t_max\": %s\n}",
                    rust_display(a.smag_alpha).c_str(), rust_display(a.eta_test).c_str(),
                    rust_display(a.nu_sm_min).c_str(), st[0], st[1], st[2], st[3],
                    rust_exp(nsum / (double)nut_m.size(), 6).c_str(), rust_exp(nmax, 6).c_str());
      if (!write_text(a.out + "/dynamic_sgs.json", dj))
        std::fprintf(stderr, "Warning : dynamic_sgs.json non écrit\n");
      std::printf("LES dynamique : cs_dyn_mean = %.4f | cs_dyn_max = %.4f (clamp [0, 0.04])\n",
                  st[0], st[1]);
    }
    std::printf("--- Résultats (champs moyens) ---\n");
    std::printf("Nuc = %.4f | Nuf = %.4f\n", nuc, nuf);
    std::printf("w_max(z=0.5) = %.5f en x = %.5f | w_max(z=0.75) = %.5f en x = %.5f\n", wm.first,
                wm.second, w75.first, w75.second);
    std::printf("y+ (paroi chaude) = %.3f | S^(1/2) = %.4f\n", dg.y_plus_hot, s_half);
    std::printf("Artefacts écrits dans %s\n", a.out.c_str());
  } catch (const std::invalid_argument& e) {
    std::fprintf(stderr, "Erreur : %s\n", e.what());
    return 1;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "Erreur : %s\n", e.what());
    return 1;
  }
  return 0;
}
