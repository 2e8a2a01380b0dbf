// clv_config.h -- run-time ingestion of the reference's two inputs:
//   Parameters.h  the #define block the reference compiles in (Parameters.h:1-53),
//                 read here as text so a run needs no recompile;
//   params.txt    key=value lines mu, P, run, seed, seed2 (main.cu:38-51).
// Host-only, header-only C++17.
#pragma once

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <string>

namespace clvhost {

// read_params (main.cu:38-51): "key=value" per line, value parsed as double.
inline std::map<std::string, double> read_params(const std::string &filename) {
  std::map<std::string, double> params;
  std::ifstream infile(filename);
  std::string line;
  while (std::getline(infile, line)) {
    std::istringstream iss(line);
    std::string key;
    double value;
    if (std::getline(iss, key, '=') && iss >> value) params[key] = value;
  }
  return params;
}

// `#define NAME value` lines of a Parameters.h-style file.  Values are kept
// as text; numeric ones convert with strtod.  Comments after the value are
// dropped.
inline std::map<std::string, std::string> read_defines(const std::string &filename) {
  std::map<std::string, std::string> out;
  std::ifstream in(filename);
  std::string line;
  while (std::getline(in, line)) {
    const size_t c = line.find("//");
    if (c != std::string::npos) line = line.substr(0, c);
    std::istringstream iss(line);
    std::string tok, name, value;
    if (!(iss >> tok) || tok != "#define") continue;
    if (!(iss >> name)) continue;
    if (!(iss >> value)) continue;
    out[name] = value;
  }
  return out;
}

// The compile-time configuration of the reference (Parameters.h:6-42), plus
// the engine's extensions (C1, KIN_DIMS, PRECISION, INIT, INIT_N0, INIT_VAR,
// CONJ_DRIFT_MODE, DT_DEVICE), all with defaults that reproduce the reference.
struct RunConfig {
  int components = 4;
  int X = 256, Y = 1, Z = 1, Nt = 32;
  double EPS = 0.01, a_x = 1.0, G = 0.1, M = 0.0, SIGMA = 0.0, D = 0.27735;
  double OMEGAX = 0.0, OMEGAY = 0.0, OMEGAZ = 0.0, Q = 0.0;
  long long STEPS = 8000000;
  double DT = 0.002;
  long long SNAPSHOT_TRIGGER = 500000, EVALUATE_TRIGGER = 100;
  // extensions
  double C1 = 0.0;
  int KIN_DIMS = 1;
  int PRECISION = 0;  // 0 FP64, 1 FP32
  int CONJ_DRIFT_MODE = 0;
  std::string INIT = "zero";  // zero | mean_field | gaussian
  double INIT_N0 = 1.0, INIT_VAR = 0.1;
  int DT_DEVICE = 0;  // 0: host-driven dt per step (reference); 1: device-resident dt
};

inline RunConfig config_from_defines(const std::map<std::string, std::string> &d) {
  RunConfig c;
  auto num = [&](const char *k, double &dst) {
    auto it = d.find(k);
    if (it != d.end()) dst = std::strtod(it->second.c_str(), nullptr);
  };
  auto inum = [&](const char *k, int &dst) {
    auto it = d.find(k);
    if (it != d.end()) dst = int(std::strtol(it->second.c_str(), nullptr, 10));
  };
  auto lnum = [&](const char *k, long long &dst) {
    auto it = d.find(k);
    if (it != d.end()) dst = std::strtoll(it->second.c_str(), nullptr, 10);
  };
  inum("COMPONENTS", c.components);
  inum("XSIZE", c.X);
  inum("YSIZE", c.Y);
  inum("ZSIZE", c.Z);
  inum("TAUSIZE", c.Nt);
  num("EPS", c.EPS);
  num("a_x", c.a_x);
  num("G", c.G);
  num("M", c.M);
  num("SIGMA", c.SIGMA);
  num("D", c.D);
  num("OMEGAX", c.OMEGAX);
  num("OMEGAY", c.OMEGAY);
  num("OMEGAZ", c.OMEGAZ);
  num("Q", c.Q);
  lnum("STEPS", c.STEPS);
  num("DT", c.DT);
  lnum("SNAPSHOT_TRIGGER", c.SNAPSHOT_TRIGGER);
  lnum("EVALUATE_TRIGGER", c.EVALUATE_TRIGGER);
  num("C1", c.C1);
  inum("KIN_DIMS", c.KIN_DIMS);
  inum("PRECISION", c.PRECISION);
  inum("CONJ_DRIFT_MODE", c.CONJ_DRIFT_MODE);
  auto it = d.find("INIT");
  if (it != d.end()) c.INIT = it->second;
  num("INIT_N0", c.INIT_N0);
  num("INIT_VAR", c.INIT_VAR);
  inum("DT_DEVICE", c.DT_DEVICE);
  return c;
}

}  // namespace clvhost
