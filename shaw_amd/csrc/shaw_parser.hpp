// Input-file parser of the drop-in shawExe: the reference's mixin parser
// (src/shared/parser/input_parser.hpp:52-70, composed at types.hpp:64-69)
// with the same sections, keys, defaults, validation messages and stdout.
// Extensions: material kind "layered" (radius/density/velocity node lists)
// rom.precision ("fp64" | "bf16x3" | "bf16x4") and rom.projection
// ("galerkin", the reference's, | "energy").
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "shaw_host.hpp"

namespace shaw {

struct InputParser {
  // general (parser_general_section.hpp:52-125)
  std::string meshDir = "empty";
  bool checkNumDispersion = true;
  bool checkCflFlag = true;
  double dt = 0, finalTime = 0;
  uint64_t nSteps = 0;
  bool exploitForcingSparsity = true;

  // io (parser_io_section.hpp:52-235)
  bool enableSnapMatrix = false, snapBinary = true;
  uint64_t vpSnapFreq = 0, spSnapFreq = 0;
  std::string vpSnapFileName = "snaps_vp", spSnapFileName = "snaps_sp";
  bool enableSeismo = false, seismoBinary = true;
  uint64_t seismoFreq = 0;
  std::string seismoFileName = "seismogram";
  std::vector<double> receivers = {5, 30, 60, 90, 120, 150, 175};

  // material (parser_material_model.hpp:48-145)
  Material material;
  std::string materialKindName;

  // source (parser_forcing_section.hpp:52-275)
  bool multiForcing = false, rank2 = false;
  int forcingSize = 1;
  SignalKind signalKind = SignalKind::Unknown;
  std::vector<double> depths, angles, periods, delays;

  // rom (parser_rom_section.hpp:52-165)
  bool romOn = false, dummyBasis = false, disableCompRomJacs = false;
  uint64_t vpRomSize = 0, spRomSize = 0;
  std::string vpBasisFile = "empty", spBasisFile = "empty";
  bool vpBasisBinary = true, spBasisBinary = true;
  int romPrecision = 0;  // SHAW_ROM_*: 0 fp64, 1 bf16x3, 2 bf16x4
  bool romEnergy = false;  // energy-consistent projection of the given bases

  // Reads `inputFile` (argc must be 2 as in input_parser.hpp:57-60).
  InputParser(int argc, char** argv);
  explicit InputParser(const std::string& inputFile) { parseAll(inputFile); }

 private:
  void parseAll(const std::string& inputFile);
};

}  // namespace shaw
