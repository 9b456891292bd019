#include "Mesh.hpp"

#include <stdexcept>
#include <string>

PointBuilder Point::create(int nsp) { return PointBuilder(nsp); }

// The checks and messages of the reference's builder (src/Mesh/PointBuilder.cpp:8-70): a
// field set twice is a logic_error; a negative temperature, a species vector of the wrong
// length, a mass fraction above 1 or a sum outside (0.99, 1.01) is a runtime_error.
PointBuilder& PointBuilder::addTemperature(double T) {
  if (p_.has_T_) throw std::logic_error("Temperature already set in the build process");
  if (T < 0) throw std::runtime_error("Temperature cannot be negative");
  p_.T_ = T;
  p_.has_T_ = true;
  return *this;
}

PointBuilder& PointBuilder::addEnthalpy(double h) {
  if (p_.has_h_) throw std::logic_error("Enthalpy already set in the build process");
  p_.h_ = h;
  p_.has_h_ = true;
  return *this;
}

PointBuilder& PointBuilder::addSpecies(std::vector<double> Y) {
  if (p_.has_Y_) throw std::logic_error("Species vector already set in the build process");
  if (static_cast<int>(Y.size()) != p_.nsp_)
    throw std::runtime_error("Vector size does not match the number of the species of the Point");
  double sum = 0.0;
  bool at_most_one = true;
  for (double v : Y) {
    sum += v;
    at_most_one = at_most_one && v <= 1.0;
  }
  if (!at_most_one || !(sum > 0.99 && sum < 1.01))
    throw std::runtime_error("Species vector breaks mass conservation law");
  p_.Y_ = std::move(Y);
  p_.has_Y_ = true;
  return *this;
}

PointBuilder& PointBuilder::addCoordinates(Coords c) {
  if (p_.has_xyz_) throw std::logic_error("Coordinates already set in the build process");
  p_.xyz_ = c;
  p_.has_xyz_ = true;
  return *this;
}

std::vector<double>& Mesh::getEnthalpyVector() {
  if (!has_h_) throw std::runtime_error("Mesh does not have enthalpy");
  return enthalpy_;
}

std::vector<Coords>& Mesh::getCoordinatesVector() {
  if (!has_xyz_) throw std::runtime_error("Mesh does not have coordinates");
  return coords_;
}

void Mesh::addPoint(const Point& p) {
  if (temperature_.empty() && p.isReady()) {
    nsp_ = p.numSpecies();
    has_h_ = p.hasEnthalpy();
    has_xyz_ = p.hasCoordinates();
  }
  const bool compatible = p.isReady() && p.numSpecies() == nsp_ && p.hasEnthalpy() == has_h_ &&
                          p.hasCoordinates() == has_xyz_;
  if (!compatible) throw std::runtime_error("Trying to add an incompatible Point to Mesh");
  temperature_.push_back(p.getTemperature());
  species_.push_back(p.getSpecies());
  if (has_h_) enthalpy_.push_back(p.getEnthalpy());
  if (has_xyz_) coords_.push_back(p.getCoordinates());
}

void Mesh::clear() {
  temperature_.clear();
  enthalpy_.clear();
  species_.clear();
  coords_.clear();
  nsp_ = 0;
  has_h_ = has_xyz_ = false;
}
