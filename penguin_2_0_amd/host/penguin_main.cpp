// penguin_main: the reference's run loop (kernel(), main.cu:113-437) as a C++ caller of the
// C ABI in include/penguin.h -- what a maintainer's main.cu becomes once the CUDA calls are
// replaced.  It builds the disk on the host (init.cpp), writes the para and initial files,
// primes the run (or resumes it from a picture), advances with the device-side landing on
// picture and end times, saves a raw snapshot at each picture, and writes one line of the
// time file and the console report every nprin steps, as the reference does.
//
//   penguin_main --out DIR [--imax N --jmax N --kmax N] [--fp32] [--eager] [--seed S]
//                [--noise A] [--ncycend N] [--nprin N] [--restart NPIC] [--set name=value ...]
//
// Files, as main.cu:118-121 names them under DIR: read/para_<label>.dat,
// files/initial_<label>.dat, files/time_<label>.dat, binary/binary_<label>_<NNNNN>.
#include <sys/stat.h>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "penguin.h"

namespace {

void usage(FILE* f) {
  std::fprintf(f,
               "usage: penguin_main --out DIR [--imax N] [--jmax N] [--kmax N] [--fp32] [--eager]\n"
               "                    [--seed S] [--noise A] [--ncycend N] [--nprin N] [--restart NPIC]\n"
               "                    [--set name=value ...]\n"
               "  --set takes any pn_params field (ndim, eos, plnt_flag, M_p, endtime, tmovie, ...)\n");
}

// one pn_params field by name: the integer flags and the double constants of global.h
bool set_param(pn_params& p, const std::string& kv) {
  const size_t eq = kv.find('=');
  if (eq == std::string::npos) return false;
  const std::string k = kv.substr(0, eq);
  const char* v = kv.c_str() + eq + 1;
  char* end = nullptr;
  const double d = std::strtod(v, &end);
  if (end == v || *end) return false;
  struct I { const char* n; int32_t* f; } ints[] = {
      {"ndim", &p.ndim},         {"eos", &p.eos},           {"flat_flag", &p.flat_flag},
      {"plnt_flag", &p.plnt_flag}, {"frrot_flag", &p.frrot_flag}, {"visc_flag", &p.visc_flag},
      {"fargo_flag", &p.fargo_flag}, {"ngeomx", &p.ngeomx},   {"ngeomy", &p.ngeomy},
      {"ngeomz", &p.ngeomz},     {"grspx", &p.grspx},       {"grspy", &p.grspy},
      {"grspz", &p.grspz},       {"init_p_iso", &p.init_p_iso}, {"nlft", &p.nlft},
      {"nrgh", &p.nrgh},         {"kill_flag", &p.kill_flag}, {"bary_flag", &p.bary_flag}};
  struct D { const char* n; double* f; } dbls[] = {
      {"p_alpha", &p.p_alpha}, {"p_beta", &p.p_beta}, {"ss_alpha", &p.ss_alpha}, {"sc_h", &p.sc_h},
      {"vis_nu", &p.vis_nu},   {"M_p", &p.M_p},       {"R_p", &p.R_p},           {"t_growth", &p.t_growth},
      {"rs_fac", &p.rs_fac},   {"gam", &p.gam},       {"courant", &p.courant},   {"endtime", &p.endtime},
      {"tmovie", &p.tmovie},   {"xmin", &p.xmin},     {"xmax", &p.xmax},         {"ymin", &p.ymin},
      {"ymax", &p.ymax},       {"zmin", &p.zmin},     {"zmax", &p.zmax}};
  for (const I& f : ints)
    if (k == f.n) {
      if (d != std::floor(d)) return false;
      *f.f = static_cast<int32_t>(d);
      return true;
    }
  for (const D& f : dbls)
    if (k == f.n) {
      *f.f = d;
      return true;
    }
  return false;
}

// get_L, get_E, get_ecc (planet.cu:11-26): the planet's orbit in the inertial frame
double orbit_L(const pn_body& b, double FrRot) { return (b.vy + FrRot) * b.x * b.x; }
double orbit_E(const pn_body& b, double FrRot) {
  const double vy = b.vy + FrRot;
  return (b.vx * b.vx + vy * vy * b.x * b.x) / 2.0 - 1.0 / b.x;
}
double orbit_ecc(const pn_body& b, double FrRot) {
  const double L = orbit_L(b, FrRot);
  return std::sqrt(1.0 + 2.0 * orbit_E(b, FrRot) * L * L);
}

}  // namespace

int main(int argc, char** argv) {
  pn_params p;
  pn_params_default(&p);
  pn_grid g{256, 768, 1};
  std::string out;
  int precision = PN_FP64, graph = 1, restart = -1;
  uint64_t seed = 1;
  double noise = 1e-4;
  long long ncycend = -1;   // default: no step limit (the run ends at endtime)
  int nprin = 100;          // global.h:70
  for (int a = 1; a < argc; ++a) {
    const std::string s = argv[a];
    auto next = [&]() -> const char* {
      if (a + 1 >= argc) {
        usage(stderr);
        std::exit(2);
      }
      return argv[++a];
    };
    if (s == "--help" || s == "-h") {
      usage(stdout);
      return 0;
    } else if (s == "--out") out = next();
    else if (s == "--imax") g.imax = std::atoi(next());
    else if (s == "--jmax") g.jmax = std::atoi(next());
    else if (s == "--kmax") g.kmax = std::atoi(next());
    else if (s == "--fp32") precision = PN_FP32;
    else if (s == "--eager") graph = 0;
    else if (s == "--seed") seed = std::strtoull(next(), nullptr, 10);
    else if (s == "--noise") noise = std::atof(next());
    else if (s == "--ncycend") ncycend = std::atoll(next());
    else if (s == "--nprin") nprin = std::atoi(next());
    else if (s == "--restart") restart = std::atoi(next());
    else if (s == "--set") {
      const char* kv = next();
      if (!set_param(p, kv)) {
        std::fprintf(stderr, "penguin_main: cannot set '%s'\n", kv);
        return 2;
      }
    } else {
      usage(stderr);
      return 2;
    }
  }
  if (out.empty() || nprin < 1) {
    usage(stderr);
    return 2;
  }

  pn_ctx* ctx = nullptr;
#define PN_CHECK(x)                                                                          \
  do {                                                                                       \
    const int rc_ = (x);                                                                     \
    if (rc_) {                                                                               \
      std::fprintf(stderr, "penguin_main: %s: %s\n", #x, ctx ? pn_last_error(ctx) : "failed"); \
      if (ctx) pn_destroy(ctx);                                                              \
      return 1;                                                                              \
    }                                                                                        \
  } while (0)

  // main.cu:113-121: the label and the files it names
  char label[512];
  PN_CHECK(pn_label(&g, &p, label, sizeof label));
  for (const char* d : {"", "/files", "/read", "/binary"}) mkdir((out + d).c_str(), 0755);
  const std::string ifname = out + "/files/initial_" + label + ".dat";
  const std::string tfname = out + "/files/time_" + label + ".dat";
  const std::string pfname = out + "/read/para_" + label + ".dat";
  const std::string bfname = out + "/binary/binary_" + label + "_";
  auto frame_path = [&](int npic) {
    char num[16];
    if (pn_frame_name(npic, num, sizeof num)) return std::string();
    return bfname + num;
  };

  // init_cells / init_bound (main.cu:277-285): the host disk and its ghost rings
  const size_t rb = pn_ring_bytes(g.jmax);
  const size_t nring = static_cast<size_t>(g.imax) * g.kmax;
  std::vector<unsigned char> rings(rb * nring), lft(rb * PN_NPAD * g.kmax), rgh(lft.size());
  std::vector<unsigned char> udr, top;
  if (p.ndim == 3) {
    udr.resize(rb * PN_NPAD * g.imax);
    top.resize(udr.size());
  }
  std::vector<pn_symdisk> val(nring);
  PN_CHECK(pn_init_disk(&g, &p, seed, noise, rings.data(), lft.data(), rgh.data(), udr.empty() ? nullptr : udr.data(),
                        top.empty() ? nullptr : top.data(), val.data()));
  PN_CHECK(pn_create(&g, &p, precision, &ctx));
  PN_CHECK(pn_load_rings(ctx, rings.data(), rings.size(), lft.data(), rgh.data(), lft.size(),
                         udr.empty() ? nullptr : udr.data(), top.empty() ? nullptr : top.data(), udr.size(),
                         val.data(), val.size()));

  // main.cu:244-262, 305-330: the para file, picture startat and the primed step state, or
  // the resumed picture
  pn_run_state st;
  int npic = 0;
  if (restart >= 0) {
    npic = restart;
    PN_CHECK(pn_load_snapshot(ctx, frame_path(npic).c_str()));
    PN_CHECK(pn_run_restart(ctx, npic, &st));
    std::printf(" Restarting from t=%g\n", st.simtime);
  } else {
    PN_CHECK(pn_write_para(&g, &p, out.c_str(), label, pfname.c_str()));
    PN_CHECK(pn_save_snapshot(ctx, frame_path(npic).c_str()));
    PN_CHECK(pn_run_begin(ctx, &st));
  }
  PN_CHECK(pn_write_initial(&g, &p, ifname.c_str()));
  FILE* tfile = std::fopen(tfname.c_str(), restart >= 0 ? "a" : "w");
  if (!tfile) {
    std::fprintf(stderr, "penguin_main: cannot open %s\n", tfname.c_str());
    pn_destroy(ctx);
    return 1;
  }

  // main.cu:335-437.  pn_run applies the landing of main.cu:352-366 on the device and
  // returns after a step that landed; the loop asks it for the steps left before the next
  // report (every nprin steps) or the step limit (the reference runs ncycend + 1 steps).
  st.land = 1;
  const double simtime0 = st.simtime;
  const long long total = ncycend >= 0 ? ncycend + 1 : -1;
  long long ncycp = 0, steps = 0;
  const auto begin = std::chrono::steady_clock::now();
  bool more = true;
  while (more) {
    long long n = nprin - ncycp;
    if (total >= 0 && total - steps < n) n = total - steps;
    const long long before = st.ncycle;
    PN_CHECK(pn_run(ctx, static_cast<int>(n), graph, &st));
    const long long done = st.ncycle - before;
    steps += done;
    ncycp += done;
    bool prin_info = false;
    if (ncycp == nprin) {
      ncycp = 0;
      prin_info = true;
    }
    if (st.event >= 1) {   // a picture time or the end time: a snapshot (main.cu:388-400)
      ++npic;
      PN_CHECK(pn_save_snapshot(ctx, frame_path(npic).c_str()));
      std::printf(" check point %d saved at %s\n", npic, frame_path(npic).c_str());
    }
    if (st.event == 2 || (total >= 0 && steps >= total)) {
      prin_info = true;
      more = false;
    }
    if (prin_info) {   // main.cu:406-436
      double mass = 0.0, torque = 0.0;
      PN_CHECK(pn_reduce(ctx, &st.planet, 0.0, 0, &mass));
      std::fprintf(tfile, "%.17g %.17g", st.simtime, mass);
      if (p.plnt_flag > 0) {
        PN_CHECK(pn_reduce(ctx, &st.planet, 0.0, 9, &torque));
        const double norm = p.M_p * p.M_p * std::pow(p.R_p, 1.0 - p.p_alpha + (p.p_beta - 1.0)) / (p.sc_h * p.sc_h);
        const pn_body& b = st.planet;
        std::fprintf(tfile, " %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g", torque / norm, b.x, b.y, b.vx,
                     b.vy, b.fx, b.fy, orbit_ecc(b, st.FrRot), -1.0 / (2.0 * orbit_E(b, st.FrRot)));
      }
      std::fprintf(tfile, "\n");
      const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - begin).count();
      const double speed = steps > 0 ? secs / static_cast<double>(steps) : 0.0;
      std::printf("\n# of steps = %lld ( t/T = %g, dt = %g )\n", static_cast<long long>(st.ncycle),
                  st.simtime / p.endtime, st.dt);
      std::printf("p1 located at r = %g and phi = %g\n", st.planet.x, st.planet.y / (2.0 * M_PI));
      std::printf("Average Speed is %g seconds per time step.\n", speed);
      if (st.simtime > simtime0)
        std::printf("Estimated time for completion: %g of %g minutes.\n\n", secs / 60.0,
                    (p.endtime - simtime0) * secs / (st.simtime - simtime0) / 60.0);
    }
  }
  std::fclose(tfile);
  return pn_destroy(ctx) == PN_OK ? 0 : 1;
}
