// gvh_config.cpp -- readers of the positional rhic-conf .properties files.
//
// The reference reads each file line by line, strips all whitespace, keeps
// what follows the first '=' (the whole line when there is none) and
// converts it with atoi/atof; keys are never compared, so only the line
// order matters (LatticeParameters.cpp:65-138, HydroParameters.cpp:49-98,
// InitialConditionParameters.cpp:65-149).  When a file has fewer lines than
// fields, std::getline fails without clearing the string, so the missing
// fields repeat the last value read; that is reproduced here.
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>

#include "../../include/gpuvh.h"
#include "gvh_internal.h"

namespace {

class PositionalReader {
 public:
  explicit PositionalReader(const std::string &path) : in_(path) {}
  bool ok() const { return in_.is_open(); }
  // next value string with the reference's stripping rules
  // next value string with the reference's stripping rules; getline into the
  // same string, so a read past the end leaves it as the stream library does
  // (unchanged when the stream was already at end of file)
  const std::string &next() {
    std::getline(in_, cur_);
    std::string s;
    for (char ch : cur_)
      if (!std::isspace((unsigned char)ch)) s.push_back(ch);
    const size_t pos = s.find('=');
    cur_ = (pos == std::string::npos) ? s : s.substr(pos + 1);
    return cur_;
  }
  int next_int() { return std::atoi(next().c_str()); }
  double next_double() { return std::atof(next().c_str()); }

 private:
  std::ifstream in_;
  std::string cur_;
};

std::string join(const char *dir, const char *name) {
  std::string d = dir ? dir : ".";
  return d + "/" + name;
}

}  // namespace

extern "C" int gvh_read_lattice_parameters(const char *config_dir, gvh_lattice_parameters *out) {
  if (!out) return gvh_fail(GVH_EINVAL, "gvh_read_lattice_parameters: out is NULL");
  const std::string path = join(config_dir, "lattice.properties");
  PositionalReader r(path);
  if (!r.ok()) return gvh_fail(GVH_EIO, "no configuration file %s", path.c_str());
  gvh_lattice_parameters L{};
  L.numLatticePointsX = r.next_int();
  L.numLatticePointsY = r.next_int();
  L.numLatticePointsRapidity = r.next_int();
  L.numProperTimePoints = r.next_int();
  L.latticeSpacingX = r.next_double();
  L.latticeSpacingY = r.next_double();
  L.latticeSpacingRapidity = r.next_double();
  L.latticeSpacingProperTime = r.next_double();
  L.numComputationalLatticePointsX = L.numLatticePointsX + 4;  // N_GHOST_CELLS
  L.numComputationalLatticePointsY = L.numLatticePointsY + 4;
  L.numComputationalLatticePointsRapidity = L.numLatticePointsRapidity + 4;
  *out = L;
  return GVH_OK;
}

extern "C" int gvh_read_hydro_parameters(const char *config_dir, gvh_hydro_parameters *out) {
  if (!out) return gvh_fail(GVH_EINVAL, "gvh_read_hydro_parameters: out is NULL");
  const std::string path = join(config_dir, "hydro.properties");
  PositionalReader r(path);
  if (!r.ok()) return gvh_fail(GVH_EIO, "no configuration file %s", path.c_str());
  gvh_hydro_parameters H{};
  H.initialProperTimePoint = r.next_double();
  H.shearViscosityToEntropyDensity = r.next_double();
  H.freezeoutTemperatureGeV = r.next_double();
  H.initializePimunuNavierStokes = r.next_int();
  H.regulationTemperatureGeV = r.next_double();
  *out = H;
  return GVH_OK;
}

extern "C" int gvh_read_ic_parameters(const char *config_dir, gvh_ic_parameters *out) {
  if (!out) return gvh_fail(GVH_EINVAL, "gvh_read_ic_parameters: out is NULL");
  const std::string path = join(config_dir, "ic.properties");
  PositionalReader r(path);
  if (!r.ok()) return gvh_fail(GVH_EIO, "no configuration file %s", path.c_str());
  gvh_ic_parameters I{};
  I.initialConditionType = r.next_int();
  I.numberOfSourceFiles = r.next_int();
  I.numberOfNucleonsPerNuclei = r.next_int();
  I.initialEnergyDensity = r.next_double();
  I.scatteringCrossSectionNN = r.next_double();
  I.impactParameter = r.next_double();
  I.fractionOfBinaryCollisions = r.next_double();
  I.rapidityVariance = r.next_double();
  I.rapidityMean = r.next_double();
  I.sourceType = r.next_int();
  *out = I;
  return GVH_OK;
}
