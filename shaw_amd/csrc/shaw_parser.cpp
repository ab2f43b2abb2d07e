#include "shaw_parser.hpp"

#include <iostream>
#include <stdexcept>

#include "yaml_lite.hpp"

namespace shaw {

namespace {

std::vector<double> scalarOrList(const yaml::Node& n) {
  if (n.isScalar()) return {n.asDouble()};
  return n.asDoubleVector();
}

void coeffs3(const yaml::Node& n, double* out, const char* what) {
  const std::vector<double> v = n.asDoubleVector();
  if (v.size() > 3) throw std::runtime_error(std::string(what) + ": at most 3 polynomial coefficients");
  for (int i = 0; i < 3; ++i) out[i] = i < (int)v.size() ? v[i] : 0.0;
}

}  // namespace

InputParser::InputParser(int argc, char** argv) {
  if (argc != 2) throw std::runtime_error("Wrong # of cmd line args, should be: ./exe <path-to-inputfile>");
  std::cout << "*** Parsing input file ***\n";
  parseAll(argv[1]);
}

void InputParser::parseAll(const std::string& inputFile) {
  const yaml::Node root = yaml::loadFile(inputFile);

  // ---- general (parser_general_section.hpp:78-125)
  const yaml::Node& g = root["general"];
  if (!g) throw std::runtime_error("General section in yaml input is mandatory!");
  if (g["meshDir"]) meshDir = g["meshDir"].str();
  else throw std::runtime_error("Empty meshDir");
  if (g["checkNumericalDispersion"]) checkNumDispersion = g["checkNumericalDispersion"].asBool();
  if (g["checkCfl"]) checkCflFlag = g["checkCfl"].asBool();
  if (g["dt"]) dt = g["dt"].asDouble();
  else throw std::runtime_error("Empty dt");
  if (g["finalTime"]) finalTime = g["finalTime"].asDouble();
  else throw std::runtime_error("Empty final time");
  if (g["exploitForcingSparsity"]) exploitForcingSparsity = g["exploitForcingSparsity"].asBool();
  if (dt <= 0.) throw std::runtime_error("Cannot have dt <= 0");
  if (finalTime <= 0.) throw std::runtime_error("Cannot have finalT <= 0");
  nSteps = static_cast<uint64_t>(finalTime / dt);
  std::cout << "\nmeshDir = " << meshDir << " \n"
            << "checkDispersion = " << std::boolalpha << checkNumDispersion << " \n"
            << "timeStep = " << dt << " \n"
            << "finalT = " << finalTime << " \n"
            << "numSteps = " << nSteps << " \n"
            << "exploitForcingSparsity " << exploitForcingSparsity << " \n";

  // ---- io (parser_io_section.hpp:110-217)
  if (const yaml::Node& io = root["io"]) {
    if (const yaml::Node& sm = io["snapshotMatrix"]) {
      enableSnapMatrix = true;
      if (sm["binary"]) snapBinary = sm["binary"].asBool();
      const yaml::Node& v = sm["velocity"];
      if (!v) throw std::runtime_error("Cannot find io inputs for velocity");
      if (v["freq"]) vpSnapFreq = v["freq"].asSize();
      else throw std::runtime_error("You must set snapshot freq for velocity");
      if (v["fileName"]) vpSnapFileName = v["fileName"].str();
      const yaml::Node& s = sm["stress"];
      if (!s) throw std::runtime_error("Cannot find io inputs for stress");
      if (s["freq"]) spSnapFreq = s["freq"].asSize();
      else throw std::runtime_error("You must set snapshot freq for stress");
      if (s["fileName"]) spSnapFileName = s["fileName"].str();
    }
    if (const yaml::Node& se = io["seismogram"]) {
      enableSeismo = true;
      if (se["binary"]) seismoBinary = se["binary"].asBool();
      if (se["fileName"]) seismoFileName = se["fileName"].str();
      if (se["freq"]) seismoFreq = se["freq"].asSize();
      else throw std::runtime_error("You must set freq for seismogram");
      if (se["receivers"] && se["receivers"].isSequence()) receivers = se["receivers"].asDoubleVector();
    }
    if (enableSnapMatrix) {
      if (vpSnapFreq <= 0) throw std::runtime_error("cannot have vpSnapshotsFreq <=0 ");
      if (spSnapFreq <= 0) throw std::runtime_error("cannot have spSnapshotsFreq <=0 ");
    }
    if (enableSeismo && seismoFreq <= 0) throw std::runtime_error("cannot have seismoFreq <=0 ");
  }
  std::cout << "\nenableSnapshotMatrix = " << std::boolalpha << enableSnapMatrix << " \n"
            << "enableSeismogram = " << enableSeismo << " \n";

  // ---- material (parser_material_model.hpp:98-145)
  const yaml::Node& m = root["material"];
  if (!m) throw std::runtime_error("Cannot find material section in yaml");
  materialKindName = m["kind"].str();
  const std::string& k = materialKindName;
  if (k == "unilayer" || k == "Unilayer" || k == "UniLayer") {
    material.kind = Material::Unilayer;
    material.params.assign(6, 0.0);
    coeffs3(m["layer"]["density"], &material.params[0], "density");
    coeffs3(m["layer"]["velocity"], &material.params[3], "velocity");
  } else if (k == "bilayer" || k == "Bilayer" || k == "BiLayer") {
    material.kind = Material::Bilayer;
    material.params.assign(13, 0.0);
    coeffs3(m["layer1"]["density"], &material.params[1], "density");
    coeffs3(m["layer1"]["velocity"], &material.params[4], "velocity");
    material.params[0] = m["layer2"]["depth"].asDouble();
    coeffs3(m["layer2"]["density"], &material.params[7], "density");
    coeffs3(m["layer2"]["velocity"], &material.params[10], "velocity");
  } else if (k == "prem" || k == "PREM") {
    material.kind = Material::Prem;
  } else if (k == "layered" || k == "Layered") {
    material.kind = Material::Layered;
    const std::vector<double> r = m["radius"].asDoubleVector();
    const std::vector<double> rho = m["density"].asDoubleVector();
    const std::vector<double> vs = m["velocity"].asDoubleVector();
    if (r.size() != rho.size() || r.size() != vs.size())
      throw std::runtime_error("layered material: radius, density and velocity lists must have equal length");
    material.params.assign(1, (double)r.size());
    material.params.insert(material.params.end(), r.begin(), r.end());
    material.params.insert(material.params.end(), rho.begin(), rho.end());
    material.params.insert(material.params.end(), vs.begin(), vs.end());
  } else if (k == "custom") {
    throw std::runtime_error(
        "material kind 'custom' is a compile-time hook of the reference executable; use kind 'layered' "
        "(radius/density/velocity node lists) for a custom radial model");
  } else {
    throw std::runtime_error("Unknown material model kind, maybe mispelled it?");
  }
  std::cout << "\nmatModelType = " << k << " \n";

  // ---- source (parser_forcing_section.hpp:93-275)
  const yaml::Node& src = root["source"];
  if (!src) throw std::runtime_error("Cannot find source section in inputfile, must be set!");
  const yaml::Node& sig = src["signal"];
  if (!sig) throw std::runtime_error("Signal node in inputfile is missing, it is mandatory!");
  if (sig["kind"]) signalKind = signalKindFromString(sig["kind"].str());
  else throw std::runtime_error("You must set the kind of the signal");
  if (sig["depth"]) depths = scalarOrList(sig["depth"]);
  else throw std::runtime_error("You must set the depth of the signal");
  if (sig["angle"]) angles = scalarOrList(sig["angle"]);
  else { angles = {0.}; std::cout << "Angle of source not specified, default==0"; }
  if (sig["period"]) periods = scalarOrList(sig["period"]);
  else throw std::runtime_error("You must set the period of the signal");
  if (sig["delay"]) delays = scalarOrList(sig["delay"]);
  else throw std::runtime_error("You must set the delay of the signal");
  if (sig["forcingSize"]) {
    forcingSize = (int)sig["forcingSize"].asInt();
    rank2 = true;
  }
  multiForcing = depths.size() >= 2 || angles.size() >= 2 || periods.size() >= 2 || delays.size() >= 2;
  if (depths.empty() || angles.empty() || periods.empty() || delays.empty())
    throw std::runtime_error("source lists must not be empty");
  if (signalKind == SignalKind::Unknown)
    throw std::runtime_error("Invalid source kind: unknown, maybe you mispelled it in input file?\n");
  for (double d : depths) if (d <= 0.) throw std::runtime_error("The depth of the source must be a positive number");
  for (double a : angles) if (a < 0.) throw std::runtime_error("The angle of source must be zero or a positive value");
  for (double p : periods) if (p <= 0.) throw std::runtime_error("Period for signal must be positive. ");
  for (double d : delays) if (d < 0.) throw std::runtime_error("Delay time of the signal must be >=0");
  if (rank2 && forcingSize < 1) throw std::runtime_error("forcingSize must be >= 1");

  // ---- rom (parser_rom_section.hpp:107-165)
  if (const yaml::Node& r = root["rom"]) {
    romOn = true;
    if (r["disableCompRomJacs"]) disableCompRomJacs = r["disableCompRomJacs"].asBool();
    const yaml::Node& v = r["velocity"];
    if (!v) throw std::runtime_error("Cannot find rom inputs for velocity");
    if (v["numModes"]) vpRomSize = v["numModes"].asSize();
    else throw std::runtime_error("You must set # of modes for velocity");
    if (v["modesFile"]) vpBasisFile = v["modesFile"].str();
    if (v["binary"]) vpBasisBinary = v["binary"].asBool();
    const yaml::Node& s = r["stress"];
    if (!s) throw std::runtime_error("Cannot find rom inputs for stress");
    if (s["numModes"]) spRomSize = s["numModes"].asSize();
    else throw std::runtime_error("You must set # of modes for stress");
    if (s["modesFile"]) spBasisFile = s["modesFile"].str();
    if (s["binary"]) spBasisBinary = s["binary"].asBool();
    if (r["precision"]) {
      const std::string p = r["precision"].str();
      if (p == "fp64") romPrecision = 0;
      else if (p == "bf16x3") romPrecision = 1;
      else if (p == "bf16x4") romPrecision = 2;
      else throw std::runtime_error("rom precision must be fp64, bf16x3 or bf16x4");
    }
    if (r["projection"]) {
      const std::string p = r["projection"].str();
      if (p == "galerkin") romEnergy = false;
      else if (p == "energy") romEnergy = true;
      else throw std::runtime_error("rom projection must be galerkin or energy");
    }
    if (vpBasisFile == "empty" || spBasisFile == "empty") {
      std::cout << "You did not set the basisfiles so I am using dummy basis" << std::endl;
      dummyBasis = true;
    }
    if (vpRomSize <= 0) throw std::runtime_error("Cannot have vpRomSize <=0 ");
    if (spRomSize <= 0) throw std::runtime_error("Cannot have spRomSize <=0 ");
  }
}

}  // namespace shaw
