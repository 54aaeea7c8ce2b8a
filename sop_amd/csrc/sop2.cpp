// sop2 -- command-file driver of libsop: one simulation, as the reference's
// sop executable runs it (sop.cu:27-417).
//
//   sop2 <input file>            run the commands of the file
//   sop2 <input file> --dry-run  parse and print the resolved run, no device
//
// Input grammar (io.cpp:9-110): whitespace-separated words; a word holding
// ';' ends a command.  Commands (sop.cu:59-83):
//   load bonds|angles|vdw|init <file>        (io.cpp:112-259)
//   set <key> <value>                         (param.cpp:233-394)
//   run                                       (simulation_ctrl, sop.cu:87-101)
// `set` keys: dynamics underdamped|overdamped, temp, nstep, istep_restart, nup,
// run, ufname, rcfname, cfname, unccfname, vfname, binfname, uncbinfname,
// cutofftype, nnlup, boxl, ncell, restart, rgen_restart, t_step, zeta, device;
// and, not in the reference: npl (pair-list interval, default 1 = every step
// as the reference) and seed (Philox key, default 1234 as sop.cu:138).
//
// Outputs per `nup` steps (sop.cu:258-272, 389-402): a line of ufname, the
// wrapped / unwrapped frames appended to binfname / uncbinfname, coordinates
// to cfname / unccfname and velocities to vfname (rewritten each time).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/sop.h"

namespace {

struct Run {
  // defaults of param.cpp:100-110 and global.cpp
  int sim_type = 1;
  float h = 2.5e-3f, halfh = 1.25e-3f, zeta = 5.0e-2f, T = 0.6f, boxl = 500.0f, ncell = 55.0f;
  double nstep = 5e7, istep_restart = 0.0;
  int nup = 1000, nnlup = 50, npl = 1, run = 1, restart = 0, rgen_restart = 0, device = 0;
  int neighborlist = 0, celllist = 0, hybridlist = 0, binsave = 1;
  unsigned long long seed = 1234;
  bool stale_halfh = false;
  std::string ufname = "update.out", rcfname = "restart_c.dat", cfname = "coord.out",
              unccfname = "unccoord.out", vfname = "veloc.out", binfname = "traj.bin",
              uncbinfname = "traj_uncorrected.bin";
  std::string f_bonds, f_angles, f_vdw, f_init;
};

[[noreturn]] void fail(const std::string &m) {
  std::fprintf(stderr, "sop2: %s\n", m.c_str());
  std::exit(1);
}

// io.cpp:9-110: commands are word sequences ended by a word holding ';'.
std::vector<std::vector<std::string>> read_input(const char *path) {
  std::ifstream in(path);
  if (!in) fail(std::string("cannot open ") + path);
  std::vector<std::vector<std::string>> cmds;
  std::vector<std::string> cur;
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream ws(line);
    std::string w;
    while (ws >> w) {
      const bool end = w.find(';') != std::string::npos;
      std::string clean;
      for (char c : w)
        if (c != ';') clean += c;
      if (!clean.empty()) cur.push_back(clean);
      if (end) {
        if (!cur.empty()) cmds.push_back(cur);
        cur.clear();
      }
    }
  }
  return cmds;
}

void set_param(Run &r, const std::vector<std::string> &c) {
  if (c.size() < 3) return;
  const std::string &k = c[1], &v = c[2];
  if (k == "dynamics") {
    if (v == "underdamped") {
      r.sim_type = 1;
      r.h = 2.5e-3f;
    } else if (v == "overdamped") {
      r.sim_type = 2;
      r.h = 0.02f;
    }
    r.halfh = (float)((double)r.h / 2.0);
  } else if (k == "temp") {
    r.T = (float)std::atof(v.c_str());
  } else if (k == "nstep") {
    r.nstep = std::atof(v.c_str());
  } else if (k == "istep_restart") {
    r.istep_restart = std::atof(v.c_str());
  } else if (k == "nup") {
    r.nup = std::atoi(v.c_str());
  } else if (k == "run") {
    r.run = std::atoi(v.c_str());
  } else if (k == "ufname") {
    r.ufname = v;
  } else if (k == "rcfname") {
    r.rcfname = v;
  } else if (k == "cfname") {
    r.cfname = v;
  } else if (k == "unccfname") {
    r.unccfname = v;
  } else if (k == "vfname") {
    r.vfname = v;
  } else if (k == "binfname") {
    r.binfname = v;
  } else if (k == "uncbinfname") {
    r.uncbinfname = v;
  } else if (k == "cutofftype") {
    r.neighborlist = v == "neighborlist";
    r.celllist = v == "celllist";
    r.hybridlist = v == "hybridlist";
  } else if (k == "nnlup") {
    r.nnlup = std::atoi(v.c_str());
  } else if (k == "boxl") {
    r.boxl = (float)std::atof(v.c_str());
  } else if (k == "ncell") {
    r.ncell = (float)std::atof(v.c_str());
  } else if (k == "restart") {
    r.restart = v == "on";
  } else if (k == "rgen_restart") {
    r.rgen_restart = v == "on";
  } else if (k == "t_step") {
    r.h = (float)std::atof(v.c_str());
    r.halfh = (float)((double)r.h / 2.0);
    r.stale_halfh = false;
  } else if (k == "zeta") {
    // param.cpp:377-386: the time step returns to the dynamics' default
    // (h/2 is not recomputed there; the engine's coefficients use the new h)
    const float h0 = r.sim_type == 1 ? 2.5e-3f : 0.02f;
    if (h0 != r.h) r.stale_halfh = true;
    r.h = h0;
    r.zeta = (float)std::atof(v.c_str());
  } else if (k == "device") {
    r.device = std::atoi(v.c_str());
  } else if (k == "npl") {
    r.npl = std::atoi(v.c_str());
  } else if (k == "seed") {
    r.seed = std::strtoull(v.c_str(), nullptr, 10);
  }
}

void load_cmd(Run &r, const std::vector<std::string> &c) {
  if (c.size() < 3) return;
  if (c[1] == "bonds") r.f_bonds = c[2];
  else if (c[1] == "angles") r.f_angles = c[2];
  else if (c[1] == "vdw") r.f_vdw = c[2];
  else if (c[1] == "init") r.f_init = c[2];
}

void print_params(const Run &r, FILE *f) {
  std::fprintf(f, "sim_type %d\nT %.6g\nh %.9g\nzeta %.9g\nboxl %.6g\nnstep %.0f\nistep_restart %.0f\n", r.sim_type,
               r.T, r.h, r.zeta, r.boxl, r.nstep, r.istep_restart);
  std::fprintf(f, "nup %d\nnnlup %d\nnpl %d\nseed %llu\nrestart %d\ndevice %d\n", r.nup, r.nnlup, r.npl, r.seed,
               r.restart, r.device);
  std::fprintf(f, "cutofftype %s\n", r.neighborlist ? "neighborlist" : r.celllist ? "celllist"
                                                     : r.hybridlist     ? "hybridlist" : "none");
  std::fprintf(f, "ufname %s\ncfname %s\nunccfname %s\nvfname %s\nbinfname %s\nuncbinfname %s\n", r.ufname.c_str(),
               r.cfname.c_str(), r.unccfname.c_str(), r.vfname.c_str(), r.binfname.c_str(), r.uncbinfname.c_str());
  std::fprintf(f, "bonds %s\nangles %s\nvdw %s\ninit %s\n", r.f_bonds.c_str(), r.f_angles.c_str(), r.f_vdw.c_str(),
               r.f_init.c_str());
}

bool load_vels(const std::string &path, int n, std::vector<float> &v) {
  std::ifstream in(path);
  if (!in) return false;
  v.assign(3 * (size_t)n, 0.0f);
  for (int i = 0; i < n; ++i) {
    int idx;
    if (!(in >> idx >> v[3 * i] >> v[3 * i + 1] >> v[3 * i + 2])) return false;
  }
  return true;
}

void save_vels(const std::string &path, int n, const std::vector<float> &v) {
  FILE *f = std::fopen(path.c_str(), "w");
  if (!f) fail("cannot write " + path);
  for (int i = 0; i < n; ++i) std::fprintf(f, "%d %f %f %f\n", i + 1, v[3 * i], v[3 * i + 1], v[3 * i + 2]);
  std::fclose(f);
}

struct Sim {
  sop_engine *e = nullptr;
  int32_t sys = -1;
  void ck(sop_status st, const char *what) const {
    if (st) fail(std::string(what) + ": " + (e ? sop_last_error(e) : "no engine"));
  }
};

// One simulation (underdamped_ctrl / overdamped_ctrl, sop.cu:104-417).
void simulate(Run &r) {
  if (r.f_bonds.empty() || r.f_vdw.empty() || r.f_init.empty()) fail("run needs load bonds, vdw and init");
  if (!r.neighborlist && !r.celllist && !r.hybridlist)
    std::fprintf(stderr, "sop2: no cutofftype set; using the neighbor / pair lists\n");
  if (r.stale_halfh)
    std::fprintf(stderr, "sop2: 'set zeta' after 't_step' reset h; coefficients use the current h\n");
  sop_topology *topo = nullptr;
  sop_status st = sop_topology_load(r.f_bonds.c_str(), r.f_angles.empty() ? nullptr : r.f_angles.c_str(),
                                    r.f_vdw.c_str(), r.f_init.c_str(), &topo);
  if (st) fail("cannot read the input files (status " + std::to_string(st) + ")");
  sop_system_desc d;
  sop_topology_desc(topo, &d);
  const int n = d.n_beads;
  sop_params &P = d.params;
  sop_params_reference(&P);
  P.T = r.T;
  P.zeta = r.zeta;
  P.h = r.h;
  P.boxl = r.boxl;
  P.n_nl = r.nnlup > 0 ? r.nnlup : 50;
  P.n_pl = r.npl > 0 ? r.npl : 1;
  P.seed = r.seed;
  P.dynamics = r.sim_type == 1 ? SOP_DYN_UNDERDAMPED : SOP_DYN_OVERDAMPED;
  std::vector<float> x(d.xyz, d.xyz + 3 * (size_t)n), v;
  double istep = 1.0;
  if (r.restart) {
    // sop.cu:157-164, 331-336: coordinates (and velocities) of a saved step
    if (sop_load_coords(r.unccfname.c_str(), n, x.data())) fail("cannot read " + r.unccfname);
    if (r.sim_type == 1 && !load_vels(r.vfname, n, v)) fail("cannot read " + r.vfname);
    istep = r.istep_restart + 1.0;
  }
  d.xyz = x.data();
  d.vel = v.empty() ? nullptr : v.data();
  d.step0 = (int64_t)istep - 1;  // Philox counter = completed steps: an exact restart
  Sim s;
  s.ck(sop_engine_create(r.device, nullptr, &s.e), "sop_engine_create");
  s.ck(sop_system_add(s.e, &d, &s.sys), "sop_system_add");
  const int n_nat = d.n_native;
  sop_topology_free(topo);

  std::FILE *out = std::fopen(r.ufname.c_str(), "a");
  if (!out) fail("cannot write " + r.ufname);
  std::vector<float> wrapped(3 * (size_t)n), unc(3 * (size_t)n), vv(3 * (size_t)n);
  auto frame = [&]() {
    s.ck(sop_get_coords(s.e, s.sys, unc.data(), nullptr), "sop_get_coords");
    s.ck(sop_get_wrapped_coords(s.e, s.sys, wrapped.data()), "sop_get_wrapped_coords");
  };
  if (r.binsave && !r.rgen_restart) {
    frame();
    sop_write_traj(r.binfname.c_str(), n, wrapped.data());
    sop_write_traj(r.uncbinfname.c_str(), n, unc.data());
  }
  const auto t0 = std::chrono::steady_clock::now();
  const int64_t first = (int64_t)istep, last = (int64_t)r.nstep;
  int64_t done = 0;
  for (int64_t k = first; k <= last;) {
    // next output step: every nup steps counted from the first step
    const int64_t next = std::min(last, k + (r.nup - 1) - ((k - first) % r.nup));
    s.ck(sop_run(s.e, &s.sys, 1, next - k + 1), "sop_run");
    done += next - k + 1;
    k = next + 1;
    if ((next - first + 1) % r.nup != 0) continue;
    double e[SOP_E_COUNT], Q = 0.0, rgsq = 0.0;
    int32_t contacts = 0;
    s.ck(sop_energy_eval(s.e, s.sys, e), "sop_energy_eval");
    s.ck(sop_observables(s.e, s.sys, &Q, &contacts, &rgsq), "sop_observables");
    double sumsq = 0.0;
    if (r.sim_type == 1) {
      s.ck(sop_get_velocities(s.e, s.sys, vv.data()), "sop_get_velocities");
      for (float c : vv) sumsq += (double)c * c;
    } else {
      s.ck(sop_get_increments(s.e, s.sys, vv.data()), "sop_get_increments");
      for (float c : vv) sumsq += (double)c * c;
      sumsq *= (double)r.zeta / (2.0 * (double)r.h);
      std::fill(vv.begin(), vv.end(), 0.0f);  // overdamped velocities stay zero (sop.cu:296-305)
    }
    const double kinT = sumsq / (3.0 * n);
    const double e_vdw_rr = e[SOP_E_NATIVE] + e[SOP_E_REP];
    if (n_nat == 0) Q = 0.0;
    if (r.sim_type == 1)
      std::fprintf(out, "%.0lf %f %f %f %f %f %f %f %f %f %d %f\n", (double)next, r.T, kinT, e[SOP_E_BOND],
                   e[SOP_E_ANGLE], e[SOP_E_NATIVE], e[SOP_E_REP], e_vdw_rr, e[SOP_E_TOTAL], Q, contacts, rgsq);
    else
      std::fprintf(out, "%.0lf %f %f %f %f %f %f %f %d %f\n", (double)next, r.T, kinT, e[SOP_E_BOND], e[SOP_E_ANGLE],
                   e_vdw_rr, e[SOP_E_TOTAL], Q, contacts, rgsq);
    std::fflush(out);
    frame();
    if (r.binsave) {
      sop_write_traj(r.binfname.c_str(), n, wrapped.data());
      sop_write_traj(r.uncbinfname.c_str(), n, unc.data());
    }
    sop_save_coords(r.cfname.c_str(), n, wrapped.data());
    sop_save_coords(r.unccfname.c_str(), n, unc.data());
    save_vels(r.vfname, n, vv);
  }
  std::fclose(out);
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::printf("Steps                  : %lld\n", (long long)done);
  std::printf("Wall Time              : %.3f sec\n", secs);
  if (done > 0) std::printf("Computation Rate       : %.3e sec / timestep\n", secs / (double)done);
  sop_engine_destroy(s.e);
}

} // namespace

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: sop2 <input file> [--dry-run]\n");
    return 2;
  }
  const bool dry = argc > 2 && std::strcmp(argv[2], "--dry-run") == 0;
  Run r;
  for (const auto &c : read_input(argv[1])) {
    if (c[0] == "load") load_cmd(r, c);
    else if (c[0] == "set") set_param(r, c);
    else if (c[0] == "run") {
      if (dry) print_params(r, stdout);
      else simulate(r);
    }
  }
  return 0;
}
