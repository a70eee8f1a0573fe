// mps_run.cpp -- command-line driver over the C ABI (include/mps.h): the
// counterpart of main.cu:3-39 and MPS::ClcMPS's time loop (kernel.cu:1872-1945).
//
//   mps_run --physical MPS_physical.txt --initial initial.txt [--out DIR]
//       reads the reference's input files (RdDat, kernel.cu:4-143) and runs
//       to FIN_TIM, writing the reference's CSV outputs every output_time.
//   mps_run --dambreak cfg0|cfg1|cfg2a|small2d|small3d [--steps N] [--every K]
//       runs a built-in dam break for N steps (output every K steps).
//   --write-reference DIR  writes the scene as MPS_physical.txt + initial.txt
//   --precision fp32|fp64, --pmax none|absg|reference, --grad standard|tensor,
//   --c C, --dt DT,
//   --graph K (replay K-step hipGraphs), --no-output, --dry-run (host only:
//   print the constants and exit).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mps.h"

namespace {

struct Scene {
  int n = 0;
  std::vector<int8_t> typ;
  std::vector<double> pos, vel, nrm;  // [3][n]
};

// MPS_physical.txt (kernel.cu:11-22): ranges, G, FIN_TIM, output_time, Ma,
// CRT_NUM, COL_RAT, DST_LMT_RAT, KNM_VSC, PCL_DST.
struct Phys {
  double rmin[3] = {0, 0, 0}, rmax[3] = {0, 0, 0}, g[3] = {0, 0, 0};
  double fin_tim = 0, output_time = 0, mach = 0.1, courant = 0.2, restitution = 0.2, dst_lmt = 0.5;
  double nu = 1e-6, l0 = 0.01;
};

struct Opts {
  std::string physical, initial, dambreak, out = "vtk_cuda", write_ref;
  int precision = -1, pmax = -1, grad = MPS_GRAD_STANDARD, steps = -1, every = 0, graph = 0;
  double c = -1, dt = -1;
  bool output = true, dry = false;
};

int fail(const char* what) {
  std::fprintf(stderr, "mps_run: %s\n", what);
  return 1;
}

int check(int rc, const char* what) {
  if (rc != MPS_OK) std::fprintf(stderr, "mps_run: %s: %s (status %d)\n", what, mps_last_error(), rc);
  return rc;
}

// The reference reads every number with fscanf("%f") into a float.
bool read_physical(const std::string& path, Phys& ph) {
  FILE* f = std::fopen(path.c_str(), "r");
  if (!f) return false;
  float v[17];
  int got = 0;
  for (int k = 0; k < 17; k++) got += std::fscanf(f, "%f", &v[k]);
  std::fclose(f);
  if (got != 17) return false;
  for (int a = 0; a < 3; a++) {
    ph.rmin[a] = v[a];
    ph.rmax[a] = v[3 + a];
    ph.g[a] = v[6 + a];
  }
  ph.fin_tim = v[9];
  ph.output_time = v[10];
  ph.mach = v[11];
  ph.courant = v[12];
  ph.restitution = v[13];
  ph.dst_lmt = v[14];
  ph.nu = v[15];
  ph.l0 = v[16];
  return true;
}

// initial.txt (kernel.cu:57-140): nP, then "index type x y z nx ny nz g" per
// particle.  Particles outside the expanded domain (MINc, MAXc) are dropped
// as RdDat does (compared in the run's precision); wall and object normals
// are negated (kernel.cu:127-128).
template <class real>
bool read_initial(const std::string& path, const Phys& ph, Scene& s, int& dropped) {
  FILE* f = std::fopen(path.c_str(), "r");
  if (!f) return false;
  int np = 0;
  if (std::fscanf(f, "%d", &np) != 1 || np < 0) {
    std::fclose(f);
    return false;
  }
  real lo[3], hi[3];
  for (int a = 0; a < 3; a++) {
    lo[a] = real(double(float(ph.rmin[a])) - 3.1 * double(float(ph.l0)));
    hi[a] = real(double(float(ph.rmax[a])) + 3.1 * double(float(ph.l0)));
  }
  std::vector<int8_t> typ;
  std::vector<double> p[3], nv[3];
  dropped = 0;
  for (int i = 0; i < np; i++) {
    int idx, t;
    float b[7];
    if (std::fscanf(f, " %d %d %f %f %f %f %f %f %f", &idx, &t, &b[0], &b[1], &b[2], &b[3], &b[4], &b[5],
                    &b[6]) != 9) {
      std::fclose(f);
      return false;
    }
    bool in = true;
    for (int a = 0; a < 3; a++) in = in && real(b[a]) < hi[a] && real(b[a]) > lo[a];
    if (!in) {
      dropped++;
      continue;
    }
    typ.push_back(int8_t(t));
    const bool wall = t == MPS_WLL || t == MPS_OBJ;
    for (int a = 0; a < 3; a++) {
      p[a].push_back(b[a]);
      nv[a].push_back(wall ? -double(b[3 + a]) : 0.0);
    }
  }
  std::fclose(f);
  s.n = (int)typ.size();
  s.typ = typ;
  s.pos.assign(3 * (size_t)s.n, 0.0);
  s.nrm.assign(3 * (size_t)s.n, 0.0);
  s.vel.assign(3 * (size_t)s.n, 0.0);
  for (int a = 0; a < 3; a++)
    for (int i = 0; i < s.n; i++) {
      s.pos[a * (size_t)s.n + i] = p[a][i];
      s.nrm[a * (size_t)s.n + i] = nv[a][i];
    }
  return true;
}

bool write_reference(const std::string& dir, const Phys& ph, const Scene& s) {
  FILE* f = std::fopen((dir + "/MPS_physical.txt").c_str(), "w");
  if (!f) return false;
  std::fprintf(f, "%.9g %.9g %.9g\n%.9g %.9g %.9g\n%.9g %.9g %.9g\n", ph.rmin[0], ph.rmin[1], ph.rmin[2],
               ph.rmax[0], ph.rmax[1], ph.rmax[2], ph.g[0], ph.g[1], ph.g[2]);
  std::fprintf(f, "%.9g\n%.9g\n%.9g\n%.9g\n%.9g\n%.9g\n%.9g\n%.9g\n", ph.fin_tim, ph.output_time, ph.mach,
               ph.courant, ph.restitution, ph.dst_lmt, ph.nu, ph.l0);
  std::fclose(f);
  f = std::fopen((dir + "/initial.txt").c_str(), "w");
  if (!f) return false;
  const size_t n = s.n;
  std::fprintf(f, "%d\n", s.n);
  for (size_t i = 0; i < n; i++)  // normals stored as in the file (RdDat negates them)
    std::fprintf(f, "%zu %d %.9g %.9g %.9g %.9g %.9g %.9g 0\n", i, int(s.typ[i]), s.pos[i], s.pos[n + i],
                 s.pos[2 * n + i], -s.nrm[i], -s.nrm[n + i], -s.nrm[2 * n + i]);
  std::fclose(f);
  return true;
}

struct Writer {
  std::string dir;
  bool on;
  FILE* open(const char* fmt, int iF) {
    char name[64];
    std::snprintf(name, sizeof(name), fmt, iF);
    std::string path = dir + "/" + name;
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) std::fprintf(stderr, "mps_run: cannot open %s\n", path.c_str());
    return f;
  }
};

struct State {
  std::vector<int8_t> typ;
  std::vector<double> pos, vel, prs, pav;
  int n = 0;
  int fetch(mps_handle* h) {
    typ.resize(n);
    pos.resize(3 * (size_t)n);
    vel.resize(3 * (size_t)n);
    prs.resize(n);
    pav.resize(n);
    return check(mps_get_state(h, typ.data(), pos.data(), vel.data(), nullptr, prs.data(), pav.data(), nullptr),
                 "mps_get_state");
  }
};

// Output_vtk(FLD, 0) + Output_vtk(MRR, 2) (kernel.cu:161-236): "%f" columns,
// pav divided by the number of steps since the last output.
int write_fluid(Writer& w, int iF, mps_handle* h, State& st, double opt_fqc) {
  if (!w.on) return MPS_OK;
  const size_t n = st.n;
  if (FILE* f = w.open("output%05d.csv", iF)) {
    std::fprintf(f, "Pos.x,Pos.y,Pos.z,Vel.x,Vel.y,Vel.z,Vel,Prs,pav\n");
    for (size_t i = 0; i < n; i++) {
      if (st.typ[i] != MPS_FLD) continue;
      const double vx = st.vel[i], vy = st.vel[n + i], vz = st.vel[2 * n + i];
      std::fprintf(f, "%f,%f,%f,%f,%f,%f,%f,%f,%f\n", st.pos[i], st.pos[n + i], st.pos[2 * n + i], vx, vy, vz,
                   std::sqrt(vx * vx + vy * vy + vz * vz), st.prs[i], st.pav[i] / opt_fqc);
    }
    std::fclose(f);
  }
  int m = 0, rc;
  if ((rc = check(mps_get_mirrors(h, nullptr, nullptr, nullptr, &m), "mps_get_mirrors"))) return rc;
  std::vector<double> mp(3 * (size_t)std::max(m, 1)), mprs(std::max(m, 1));
  if ((rc = check(mps_get_mirrors(h, nullptr, mp.data(), mprs.data(), &m), "mps_get_mirrors"))) return rc;
  if (FILE* f = w.open("outputMRR%05d.csv", iF)) {
    std::fprintf(f, "PosM.x,PosM.y,PosM.z,VelM,PrsM\n");
    for (int k = 0; k < m; k++)
      std::fprintf(f, "%f,%f,%f,%f,%f,\n", mp[k], mp[(size_t)m + k], mp[2 * (size_t)m + k], 0.0, mprs[k]);
    std::fclose(f);
  }
  return MPS_OK;
}

// Output_vtk(WLL, 1) + Output_vtk(WLL, 3) (kernel.cu:197-255).
void write_walls(Writer& w, const State& st, const Scene& s, const std::vector<int8_t>& edge) {
  if (!w.on) return;
  const size_t n = st.n;
  if (FILE* f = w.open("outputwall%05d.csv", 0)) {
    std::fprintf(f, "Pos.x,Pos.y,Pos.z,Prs,pav,SE\n");
    for (size_t i = 0; i < n; i++)
      if (st.typ[i] == MPS_WLL)
        std::fprintf(f, "%f,%f,%f,%f,%f,%d\n", st.pos[i], st.pos[n + i], st.pos[2 * n + i], st.prs[i], st.pav[i],
                     int(edge[i]));
    std::fclose(f);
  }
  if (FILE* f = w.open("outputWLLvec%05d.csv", 0)) {
    std::fprintf(f, "PosM.x,PosM.y,PosM.z,vec.x,vec.y,vec.z\n");
    for (size_t i = 0; i < n; i++)
      if (st.typ[i] == MPS_WLL)
        std::fprintf(f, "%f,%f,%f,%f,%f,%f\n", st.pos[i], st.pos[n + i], st.pos[2 * n + i], s.nrm[i],
                     s.nrm[n + i], s.nrm[2 * n + i]);
    std::fclose(f);
  }
}

// Output_vtk2 (kernel.cu:258-283): the settled fluid as an initial.txt.
void write_seitei(Writer& w, const State& st) {
  if (!w.on) return;
  const size_t n = st.n;
  if (FILE* f = w.open("seitei.txt", 0)) {
    int nf = 0;
    for (size_t i = 0; i < n; i++) nf += st.typ[i] == MPS_FLD;
    std::fprintf(f, "%d\n", nf);
    int k = 0;
    for (size_t i = 0; i < n; i++)
      if (st.typ[i] == MPS_FLD)
        std::fprintf(f, "%d %d %f %f %f %f %f %f %f\n", k++, int(st.typ[i]), st.pos[i], st.pos[n + i],
                     st.pos[2 * n + i], 0.0, 0.0, 0.0, 0.0);
    std::fclose(f);
  }
}

// WrtDat (kernel.cu:286-299): fluid + mirror output, then pav = 0.
int wrt_dat(Writer& w, int iF, mps_handle* h, State& st, double opt_fqc) {
  int rc;
  if ((rc = st.fetch(h)) || (rc = write_fluid(w, iF, h, st, opt_fqc))) return rc;
  std::vector<double> zero(st.n, 0.0);
  return check(mps_set_state(h, nullptr, nullptr, nullptr, nullptr, zero.data()), "mps_set_state");
}

// ClcMPS's clock (kernel.cu:1872-1945), with TIM, outtime and OPT_FQC in the
// run's precision so that the output instants match.  Steps between two
// outputs are issued as one mps_step call.
template <class real>
int run_clock(mps_handle* h, Writer& w, State& st, double dt_d, double output_time_d, double fin_tim_d,
              long long& steps_done) {
  const real dt = real(dt_d), output_time = real(output_time_d), fin = real(fin_tim_d);
  real TIM = 0, outtime = 0, OPT_FQC = 0;
  int iF = 1, rc;
  long long pending = 0;
  for (;;) {
    if (outtime >= output_time) {
      outtime -= output_time;
      if ((rc = check(mps_step(h, (int)pending), "mps_step"))) return rc;
      steps_done += pending;
      pending = 0;
      if ((rc = wrt_dat(w, iF, h, st, double(OPT_FQC)))) return rc;
      std::printf("Time = %f\n", double(TIM));
      iF++;
      if (TIM >= fin) {
        write_seitei(w, st);
        return MPS_OK;
      }
      OPT_FQC = 0;
    }
    pending++;
    outtime += dt;
    TIM += dt;
    OPT_FQC += real(1);
    if (pending >= (1 << 20)) {
      if ((rc = check(mps_step(h, (int)pending), "mps_step"))) return rc;
      steps_done += pending;
      pending = 0;
    }
  }
}

// Built-in dam breaks (the workloads of the tests and bench).
bool builtin(const std::string& name, mps_dambreak_spec& sp, Phys& ph, double& dt, int& prec, int& steps) {
  std::memset(&sp, 0, sizeof(sp));
  sp.wall_layers = 3;
  sp.domain_top = 1.0;
  ph.g[1] = -9.81;
  ph.nu = 1e-6;
  ph.restitution = 0.2;
  ph.dst_lmt = 0.5;
  if (name == "cfg0") {
    sp.dim = 2, sp.l0 = 0.01, sp.tank[0] = 1.0, sp.tank[1] = 0.6, sp.column[0] = 0.25, sp.column[1] = 0.5;
    dt = 5e-4, prec = MPS_FP64, steps = 2000;
  } else if (name == "cfg1" || name == "cfg2a") {
    const bool big = name == "cfg2a";
    sp.dim = 3, sp.l0 = big ? 0.0025 : 0.005, sp.seed = big ? 2 : 1, sp.jitter = 0.01;
    sp.tank[0] = 1.0, sp.tank[1] = 0.6, sp.tank[2] = 0.3;
    sp.column[0] = 0.25, sp.column[1] = 0.5, sp.column[2] = 0.3;
    dt = big ? 1.25e-4 : 2.5e-4, prec = MPS_FP32, steps = big ? 500 : 2000;
  } else if (name == "small2d") {
    sp.dim = 2, sp.l0 = 0.01, sp.tank[0] = 0.3, sp.tank[1] = 0.3, sp.column[0] = 0.1, sp.column[1] = 0.2;
    sp.domain_top = 0.5, dt = 5e-4, prec = MPS_FP64, steps = 200;
  } else if (name == "small3d") {
    sp.dim = 3, sp.l0 = 0.02, sp.seed = 5, sp.jitter = 0.01, sp.domain_top = 0.5;
    sp.tank[0] = 0.4, sp.tank[1] = 0.3, sp.tank[2] = 0.2, sp.column[0] = 0.1, sp.column[1] = 0.2, sp.column[2] = 0.1;
    dt = 1e-3, prec = MPS_FP64, steps = 100;
  } else {
    return false;
  }
  ph.l0 = sp.l0;
  ph.fin_tim = steps * dt;  // for --write-reference: 20 outputs over the run
  ph.output_time = ph.fin_tim / 20;
  return true;
}

int usage() {
  std::fprintf(stderr,
               "usage: mps_run (--physical F --initial F | --dambreak cfg0|cfg1|cfg2a|small2d|small3d)\n"
               "               [--out DIR] [--steps N] [--every K] [--precision fp32|fp64]\n"
               "               [--pmax none|absg|reference] [--grad standard|tensor]\n"
               "               [--c C] [--dt DT] [--graph K]\n"
               "               [--no-output] [--dry-run] [--write-reference DIR]\n");
  return 2;
}

}  // namespace

int main(int argc, char** argv) {
  const auto t_start = std::chrono::steady_clock::now();
  Opts o;
  for (int i = 1; i < argc; i++) {
    const std::string a = argv[i];
    auto next = [&]() -> const char* { return i + 1 < argc ? argv[++i] : nullptr; };
    const char* v = nullptr;
    if (a == "--no-output") o.output = false;
    else if (a == "--dry-run") o.dry = true;
    else if (!(v = next())) return usage();
    else if (a == "--physical") o.physical = v;
    else if (a == "--initial") o.initial = v;
    else if (a == "--dambreak") o.dambreak = v;
    else if (a == "--out") o.out = v;
    else if (a == "--write-reference") o.write_ref = v;
    else if (a == "--steps") o.steps = std::atoi(v);
    else if (a == "--every") o.every = std::atoi(v);
    else if (a == "--graph") o.graph = std::atoi(v);
    else if (a == "--c") o.c = std::atof(v);
    else if (a == "--dt") o.dt = std::atof(v);
    else if (a == "--precision") o.precision = std::string(v) == "fp64" ? MPS_FP64 : MPS_FP32;
    else if (a == "--grad") {
      if (std::string(v) == "tensor") o.grad = MPS_GRAD_TENSOR;
      else if (std::string(v) == "standard") o.grad = MPS_GRAD_STANDARD;
      else return usage();
    } else if (a == "--pmax")
      o.pmax = std::string(v) == "none" ? MPS_PMAX_NONE : std::string(v) == "absg" ? MPS_PMAX_ABS_G
                                                                                  : MPS_PMAX_REFERENCE;
    else return usage();
  }
  const bool from_files = !o.physical.empty();
  if (from_files == !o.dambreak.empty()) return usage();

  Phys ph;
  Scene s;
  mps_params p;
  mps_params_default(&p);
  int dim = 3, prec = MPS_FP32, def_steps = -1;
  double dt = -1;
  if (from_files) {
    // RdDat + SetPara: c and dt derived from Ma and umax (kernel.cu:407-415)
    if (!read_physical(o.physical, ph)) return fail("cannot read the physical-parameter file");
    prec = o.precision >= 0 ? o.precision : MPS_FP32;
    int dropped = 0;
    const bool ok = prec == MPS_FP64 ? read_initial<double>(o.initial, ph, s, dropped)
                                     : read_initial<float>(o.initial, ph, s, dropped);
    if (!ok) return fail("cannot read the initial-particle file");
    dim = ph.rmax[2] > ph.rmin[2] ? 3 : 2;
    std::printf("RdDat: %d particles (%d outside the domain dropped)\n", s.n, dropped);
    p.pmax_mode = MPS_PMAX_REFERENCE;
  } else {
    mps_dambreak_spec sp;
    if (!builtin(o.dambreak, sp, ph, dt, prec, def_steps)) return fail("unknown --dambreak workload");
    if (o.precision >= 0) prec = o.precision;
    dim = sp.dim;
    int nfl = 0, nwl = 0;
    if (check(mps_scene_dambreak_count(&sp, &nfl, &nwl), "mps_scene_dambreak_count")) return 1;
    s.n = nfl + nwl;
    s.typ.resize(s.n);
    s.pos.assign(3 * (size_t)s.n, 0.0);
    s.nrm.assign(3 * (size_t)s.n, 0.0);
    s.vel.assign(3 * (size_t)s.n, 0.0);
    if (check(mps_scene_dambreak_fill(&sp, s.typ.data(), s.pos.data(), s.nrm.data(), ph.rmin, ph.rmax),
              "mps_scene_dambreak_fill"))
      return 1;
    p.pmax_mode = MPS_PMAX_ABS_G;
    p.c = 22.0;
  }
  p.dim = dim;
  p.precision = prec;
  p.l0 = ph.l0;
  for (int a = 0; a < 3; a++) {
    p.range_min[a] = ph.rmin[a];
    p.range_max[a] = ph.rmax[a];
    p.gravity[a] = ph.g[a];
  }
  p.nu = ph.nu;
  p.mach = ph.mach;
  p.restitution = ph.restitution;
  p.dst_lmt_rat = ph.dst_lmt;
  p.dt = o.dt > 0 ? o.dt : dt;
  if (o.c > 0) p.c = o.c;
  if (o.pmax >= 0) p.pmax_mode = o.pmax;
  p.grad_model = o.grad;

  if (!o.write_ref.empty()) {
    if (!write_reference(o.write_ref, ph, s)) return fail("cannot write the reference-format files");
    std::printf("wrote %s/MPS_physical.txt and %s/initial.txt\n", o.write_ref.c_str(), o.write_ref.c_str());
  }
  mps_scene sc = {s.n, 0, s.typ.data(), s.pos.data(), s.vel.data(), s.nrm.data(), nullptr, nullptr, nullptr};
  mps_constants k;
  if (check(mps_compute_constants(&p, &sc, &k), "mps_compute_constants")) return 1;
  std::printf("SetPara: n0=%.9g n0_grad=%.9g lambda=%.9g nB=%dx%dx%d dt=%.9g c=%.9g umax=%.9g Pmax=%.9g\n",
              k.n0, k.n0_grad, k.lmd, k.nb[0], k.nb[1], k.nb[2], k.dt, k.snd, k.umax, k.pmax);
  std::printf("fluid %d, walls %d\n", k.n_fluid, k.n_wall);
  std::printf("pressure gradient: %s\n", p.grad_model == MPS_GRAD_TENSOR ? "tensor" : "standard");
  if (o.dry) return 0;

  mps_handle* h = nullptr;
  if (check(mps_create(&p, &sc, &h), "mps_create")) return 1;
  if (o.graph > 0 && check(mps_set_graph(h, o.graph), "mps_set_graph")) return 1;
  Writer w{o.out, o.output};
  State st;
  st.n = s.n;
  std::vector<int8_t> edge(s.n, 0);
  if (check(mps_compute_edge(&p, &sc, edge.data()), "mps_compute_edge")) return 1;
  // ClcMPS prologue: walls once, the initial fluid (OPT_FQC = 1)
  if (st.fetch(h)) return 1;
  write_walls(w, st, s, edge);
  if (wrt_dat(w, 0, h, st, 1.0)) return 1;

  long long steps_done = 0;
  const auto t0 = std::chrono::steady_clock::now();
  int rc;
  if (from_files && o.steps < 0) {
    rc = prec == MPS_FP64 ? run_clock<double>(h, w, st, k.dt, ph.output_time, ph.fin_tim, steps_done)
                          : run_clock<float>(h, w, st, k.dt, ph.output_time, ph.fin_tim, steps_done);
  } else {
    const int total = o.steps >= 0 ? o.steps : def_steps;
    const int every = o.every > 0 ? o.every : total;
    rc = MPS_OK;
    int iF = 1;
    for (int done = 0; done < total && rc == MPS_OK; iF++) {
      const int chunk = std::min(every, total - done);
      if ((rc = check(mps_step(h, chunk), "mps_step"))) break;
      done += chunk;
      steps_done = done;
      if ((rc = wrt_dat(w, iF, h, st, double(chunk)))) break;
      std::printf("step %d, Time = %f\n", done, done * k.dt);
    }
    if (rc == MPS_OK) write_seitei(w, st);
  }
  if (rc == MPS_OK) rc = check(mps_sync(h), "mps_sync");
  const double run_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  int live = 0;
  if (rc == MPS_OK && st.fetch(h) == MPS_OK)
    for (int i = 0; i < s.n; i++) live += st.typ[i] == MPS_FLD;
  mps_destroy(h);
  if (rc != MPS_OK) return 1;
  const double total_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  std::printf("end MPS: %lld steps, %d live fluid, loop %.3f s (with output), %.3e particle-steps/s\n", steps_done,
              live, run_s, run_s > 0 ? double(live) * double(steps_done) / run_s : 0.0);
  std::printf("%.3f second\n", total_s);
  return 0;
}
