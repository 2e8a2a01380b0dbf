// clv_main.cpp -- host driver of the engine, in the roles of the reference's
// main.cu / Lattice / Propagator / Observables / SaveResults, calling the HIP
// kernels only through the C ABI of clv.h.
//
//   clv_main [device] [--params-h PATH] [--params PATH] [--dry-run]
//
// Inputs are read at run time: Parameters.h (default ./Parameters.h) for the
// macros the reference compiles in, params.txt (default ./params.txt) for
// mu, P, run, seed, seed2 (main.cu:79-92).  Outputs are the reference's:
// ./DATA/BoseCL_Sim_mu_<MU>_P_<P>/Run<r>/{Parameters.txt, F_x.txt, F_y.txt,
// F_z.txt, rawdata/} and the shared ./DATA/BoseCL_Sim_mu_<MU>_P_<P>/Skalars.txt,
// with the reference's stdout lines.
#include <sys/stat.h>
#include <sys/types.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "clv.h"
#include "clv_config.h"

using namespace std;
using clvhost::RunConfig;

namespace {

void die(const clv_ctx *ctx, const string &what) {
  cerr << "clv error in " << what << ": " << clv_last_error(ctx) << endl;
  exit(2);
}

// One chain's psi and psibar on the device (the reference's pair of
// ComplexLattice objects `fields` / `fields_conjug`, main.cu:94-95).
class Lattice {
 public:
  Lattice(const RunConfig &c, int device, const clv_params &p) : cfg(c) {
    clv_shape s{c.X, c.Y, c.Z, c.Nt};
    if (clv_create(&s, 1, c.PRECISION ? CLV_FP32 : CLV_FP64, device, &ctx) != CLV_OK)
      die(nullptr, "clv_create");
    if (clv_set_params(ctx, 0, &p) != CLV_OK) die(ctx, "clv_set_params");
  }
  ~Lattice() { clv_destroy(ctx); }
  // set_value (Lattice.cu:50-81) on both fields: components 0 and 3 get
  // (re, im), components 1 and 2 get 2/3 of it; psibar holds the conjugate.
  void set_mean_field(double re, double im) {
    const size_t n = size_t(4) * cfg.X * cfg.Y * cfg.Z * cfg.Nt;
    vector<double> psi(2 * n), psib(2 * n);
    const size_t per = n / 4;
    for (int c = 0; c < 4; ++c) {
      const double f = (c == 0 || c == 3) ? 1.0 : 2.0 / 3.0;
      for (size_t k = 0; k < per; ++k) {
        const size_t i = c * per + k;
        psi[2 * i] = f * re;
        psi[2 * i + 1] = f * im;
        psib[2 * i] = f * re;
        psib[2 * i + 1] = -f * im;
      }
    }
    if (clv_set_fields(ctx, 0, psi.data(), psib.data()) != CLV_OK) die(ctx, "clv_set_fields");
  }
  void init_gaussian(double n0, double var) {
    if (clv_init_fields(ctx, 0, n0, var) != CLV_OK) die(ctx, "clv_init_fields");
  }
  clv_ctx *ctx = nullptr;
  RunConfig cfg;
};

// Propagator::propagate (Propagator_without_cutoff.cu:406-442): drift and the
// max/mean reduction on the device, dt on the host with the reference's
// clamp messages (:393-400), then the update with that dt.
class Propagator {
 public:
  Propagator(Lattice &l, const clv_params &p) : lat(l), prm(p) {}
  void propagate() {
    if (lat.cfg.DT_DEVICE) {
      if (clv_propagate(lat.ctx, 1, 0) != CLV_OK) die(lat.ctx, "clv_propagate");
      return;
    }
    double max_drift = 0.0, mean_drift = 0.0;
    if (clv_drift_stats(lat.ctx, 0, &max_drift, &mean_drift) != CLV_OK) die(lat.ctx, "clv_drift_stats");
    int32_t clamp = 0;
    const double dt = clv_compute_new_dt(mean_drift, max_drift, &prm, &clamp);
    if (clamp < 0) cout << "dt_new is too small, setting to min_dt: " << prm.dt_min << endl;
    if (clamp > 0) cout << "dt_new is too large, setting to max_dt: " << prm.dt_max << endl;
    if (clv_apply_update(lat.ctx, 0, dt) != CLV_OK) die(lat.ctx, "clv_apply_update");
  }

 private:
  Lattice &lat;
  clv_params prm;
};

// Observables (Observables_without_cutoff.cu:676-807): the enabled set of
// main.cu:61-77 -- n_tot, n_11..n_22, F_x, F_y, F_z -- plus the energy.
class Observables {
 public:
  explicit Observables(Lattice &l) : lat(l) {}
  void evaluate() {
    if (clv_evaluate(lat.ctx, 0) != CLV_OK) die(lat.ctx, "clv_evaluate");
  }
  void read() {
    const size_t vs = size_t(lat.cfg.X) * lat.cfg.Y * lat.cfg.Z;
    Fx.assign(vs, 0.0);
    Fy.assign(vs, 0.0);
    Fz.assign(vs, 0.0);
    if (clv_read_observables(lat.ctx, 0, &o, Fx.data(), Fy.data(), Fz.data()) != CLV_OK)
      die(lat.ctx, "clv_read_observables");
  }
  int get_counter() {
    read();
    return int(o.count);
  }
  clv_observables o{};
  vector<double> Fx, Fy, Fz;

 private:
  Lattice &lat;
};

bool folder_exists(const string &foldername) {
  struct stat st;
  if (stat(foldername.c_str(), &st) == 0) return S_ISDIR(st.st_mode) != 0;
  return false;
}

// SaveResults (SaveResults.cu:22-795), the parts main.cu enables.
class SaveResults {
 public:
  SaveResults(Observables *observables, unsigned long long seedval, const RunConfig &c)
      : obs(observables), seedvalue(seedval), cfg(c) {}

  bool make_directory(double P, double MU) {  // SaveResults.cu:52-84
    stringstream dirname;
    dirname << "./DATA/BoseCL_Sim_mu_" << MU << "_P_" << P << "/Run";
    for (int r = 1; r < 1000; r++) {
      if (!folder_exists(dirname.str() + to_string(r))) {
        dirname << r;
        break;
      }
    }
    std::filesystem::path dirPath = dirname.str();
    try {
      if (std::filesystem::create_directories(dirPath)) {
        std::cout << "Directory created successfully: " << dirPath << std::endl;
        directory = dirname.str();
        return true;
      } else {
        std::cout << "Directory already exists: " << dirPath << std::endl;
        return false;
      }
    } catch (const std::filesystem::filesystem_error &e) {
      std::cerr << "Error: " << e.what() << std::endl;
      return false;
    }
  }

  bool make_directory_rawdata() {  // SaveResults.cu:151-164
    directory_rawdata = directory + "/rawdata";
    if (mkdir(directory_rawdata.c_str(), 0755) < 0) {
      cout << "Directory for raw data could not be created!" << endl;
      return false;
    }
    cout << "Directory for raw data was successfully created" << endl;
    return true;
  }

  void save_parameters(double P, double MU) {  // SaveResults.cu:166-192
    fstream output;
    stringstream temp;
    temp << directory << "/Parameters.txt";
    output.open(temp.str(), fstream::out);
    output << "T\t" << 1. / (cfg.Nt * cfg.EPS) << endl
           << "mu\t" << MU << endl
           << "g\t" << cfg.G << endl
           << "M\t" << cfg.M << endl
           << "SIGMA\t" << cfg.SIGMA << endl
           << "D\t" << cfg.D << endl
           << "OMEGAX\t" << cfg.OMEGAX << endl
           << "OMEGAY\t" << cfg.OMEGAY << endl
           << "OMEGAZ\t" << cfg.OMEGAZ << endl
           << "P\t" << P << endl
           << "Q\t" << cfg.Q << endl
           << "components\t" << cfg.components << endl
           << "lattice\t" << cfg.X << "x" << cfg.Y << "x" << cfg.Z << "x" << cfg.Nt << endl
           << "dt\t" << cfg.DT << endl
           << "steps\t" << cfg.STEPS << endl
           << "SNAPSHOT_TRIGGER\t" << cfg.SNAPSHOT_TRIGGER << endl
           << "EVALUATE_TRIGGER\t" << cfg.EVALUATE_TRIGGER << endl
           << "seed\t" << seedvalue;
    output.close();
    cout << "Saved parameters" << endl;
  }

  void save_snapshot(double P, double MU) {  // SaveResults.cu:194-607
    const int number = obs->get_counter();
    const clv_observables &o = obs->o;
    {
      fstream output;
      stringstream directory_scalars, temp;
      directory_scalars << "./DATA/BoseCL_Sim_mu_" << MU << "_P_" << P;
      temp << directory_scalars.str() << "/Skalars.txt";
      output.open(temp.str(), fstream::app);
      const double nan = std::nan("");
      const double c = double(o.count);
      // write_n_* divide by the evaluation counter; with none yet the
      // reference divides by zero
      auto val = [&](double v) { return c > 0 ? v : nan; };
      string line;
      line += to_string(val(o.n_tot)) + "\t";
      for (int k = 0; k < 4; ++k) line += to_string(val(o.n_comp[k])) + "\t";
      line += "\t\t";  // P2 disabled (main.cu:64)
      output << line << endl;
      if (output.good()) cout << "Saved scalars at time " << number << endl;
      else cout << "Saving scalars at time " << number << " failed" << endl;
      output.close();
    }
    write_field("F_x", obs->Fx, number);
    write_field("F_y", obs->Fy, number);
    write_field("F_z", obs->Fz, number);
  }

 private:
  void write_field(const char *name, const vector<double> &F, int number) {
    fstream output;
    stringstream temp;
    temp << directory << "/" << name << ".txt";
    output.open(temp.str(), fstream::app);
    for (double v : F) output << v << "\t";
    output << endl;
    if (output.good()) cout << "Saved real-space " << name << " at time " << number << endl;
    else cout << "Saving real-space " << name << " at time " << number << " failed" << endl;
    output.close();
  }

  Observables *obs;
  unsigned long long seedvalue;
  RunConfig cfg;
  string directory, directory_rawdata;
};

}  // namespace

int main(int argc, char **argv) {
  int devnum = 0;
  string params_h = "Parameters.h", params_txt = "params.txt";
  bool dry = false;
  for (int a = 1; a < argc; ++a) {
    const string s = argv[a];
    if (s == "--params-h" && a + 1 < argc) params_h = argv[++a];
    else if (s == "--params" && a + 1 < argc) params_txt = argv[++a];
    else if (s == "--dry-run") dry = true;
    else devnum = atoi(argv[a]);
  }
  double timer = time(NULL);
  std::ifstream probe(params_h);
  if (!probe.good()) {
    cerr << "cannot read " << params_h << endl;
    return 1;
  }
  const RunConfig cfg = clvhost::config_from_defines(clvhost::read_defines(params_h));
  auto params = clvhost::read_params(params_txt);
  double MU = 0.0, P = 0.0;
  try {
    MU = params.at("mu");
    P = params.at("P");
  } catch (const std::out_of_range &e) {
    std::cerr << "Fehler: params.txt unvollständig (mu/P fehlen?)" << std::endl;
    return 1;
  }
  unsigned long long seedvalue = static_cast<unsigned long long>(params["seed"]);

  clv_params p;
  clv_default_params(&p);
  p.mu = MU;
  p.P = P;
  p.G = cfg.G;
  p.c1 = cfg.C1;
  p.Q = cfg.Q;
  p.eps = cfg.EPS;
  p.a_x = cfg.a_x;
  p.DT = cfg.DT;
  p.seed = seedvalue;
  p.kin_dims = cfg.KIN_DIMS;
  p.conj_drift_mode = cfg.CONJ_DRIFT_MODE;

  Lattice *lat = nullptr;
  Propagator *prop = nullptr;
  Observables *obs = nullptr;
  if (!dry) {
    lat = new Lattice(cfg, devnum, p);
    prop = new Propagator(*lat, p);
    obs = new Observables(*lat);
  }
  SaveResults *save = new SaveResults(obs, seedvalue, cfg);
  if (!save->make_directory(P, MU)) return 1;
  save->save_parameters(P, MU);
  save->make_directory_rawdata();

  if (!dry) {
    if (cfg.INIT == "mean_field") lat->set_mean_field(std::sqrt(std::abs(MU) / cfg.G / cfg.components), 0.);
    else if (cfg.INIT == "gaussian") lat->init_gaussian(cfg.INIT_N0, cfg.INIT_VAR);
    else lat->set_mean_field(0., 0.);
  }

  cout << "Started simulation with parameters" << endl
       << "T " << 1. / (cfg.Nt * cfg.EPS) << endl
       << "mu " << MU << endl
       << "g " << cfg.G << endl
       << "M " << cfg.M << endl
       << "SIGMA " << cfg.SIGMA << endl
       << "D " << cfg.D << endl
       << "OMEGAX " << cfg.OMEGAX << endl
       << "OMEGAY " << cfg.OMEGAY << endl
       << "OMEGAZ " << cfg.OMEGAZ << endl
       << "P " << P << endl
       << "components " << cfg.components << endl
       << "lattice " << cfg.X << "x" << cfg.Y << "x" << cfg.Z << "x" << cfg.Nt << endl
       << "dt " << cfg.DT << endl
       << "steps " << cfg.STEPS << endl
       << "seed " << seedvalue << endl;
  if (dry) return 0;

  // main loop (main.cu:144-157)
  for (long long i = 0; i < cfg.STEPS; i++) {
    const bool is_snapshot_step = ((i % cfg.SNAPSHOT_TRIGGER) == 0) && i != 0 && i != cfg.STEPS - 1;
    if (is_snapshot_step || i == 1) save->save_snapshot(P, MU);
    prop->propagate();
    if (i % cfg.EVALUATE_TRIGGER == 0) obs->evaluate();
  }
  save->save_snapshot(P, MU);
  cout << "Finished simulation after computation time of " << time(NULL) - timer << " seconds." << endl;
  delete save;
  delete obs;
  delete prop;
  delete lat;
  return 0;
}
